"""Run-file and result-file handling (wanniernlqg_amd/runio.py) on the CPU.

Pinned by the reference's own tests: the writer precision values of
test/result_writer_precision_unit.jl and the energy-axis cases of
test/task_input_energy_axes_unit.jl (restated as Python calls); the column order by a tensor whose
entries encode their own indices; the SHG rows by a values array whose totals are known."""
import os
import shutil
from fractions import Fraction

import numpy as np
import pytest

from wanniernlqg_amd import runio as R

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "synthetic_runtime", "synthetic_tb.dat")


def test_writer_precision_reference_values(tmp_path):
    tensor = np.full((1, 1, 1, 1), complex(1.0 / 3.0, -2.0 / 7.0))
    p7 = R.write_response_tensor(0.0, [0.125], tensor, 1, str(tmp_path), "default.dat")
    p15 = R.write_response_tensor(0.0, [0.125], tensor, 1, str(tmp_path), "precise.dat", digits=15)
    f7 = open(p7).read().splitlines()[2].split()
    f15 = open(p15).read().splitlines()[2].split()
    assert f7[2] == "3.3333333e-01"
    assert f15[2] == "3.333333333333333e-01"
    assert f15[3] == "-2.857142857142857e-01"
    with pytest.raises(ValueError):
        R.write_response_tensor(0.0, [0.125], tensor, 1, str(tmp_path), "invalid.dat", digits=6)
    with pytest.raises(ValueError):
        R.write_response_tensor(0.0, [0.125], tensor, 1, str(tmp_path), "invalid.dat", digits=18)


def test_writer_columns_rightmost_fastest(tmp_path):
    omega = np.array([0.5, 1.25])
    t = np.zeros((2, 3, 3, 3), dtype=np.complex128)
    for w in range(2):
        for a in range(3):
            for b in range(3):
                for c in range(3):
                    t[w, a, b, c] = complex(100 * a + 10 * b + c + 1000 * w, -(a + 1))
    path = R.write_response_tensor(0.3, omega, t, 3, str(tmp_path / "sub"), "x_sc_conv.dat", digits=17,
                                   region_metadata={"units": "A/V^2"})
    lines = open(path).read().splitlines()
    assert lines[0] == "#### WannierNLQG response tensor"
    assert lines[1] == "# units=A/V^2"
    labels = lines[2].split()[1:]
    assert labels[:2] == ["Efermi", "Omega"]
    assert labels[2:5] == ["xxx", "xxy", "xxz"] and labels[8] == "xzx" and labels[11] == "yxx" and labels[-1] == "zzz"
    assert len(labels) == 2 + 27
    for w, line in enumerate(lines[3:]):
        f = [float(x) for x in line.split()]
        assert f[0] == 0.3 and f[1] == omega[w]
        vals = np.array(f[2:]).reshape(27, 2)
        for i, (a, b, c) in enumerate(np.ndindex(3, 3, 3)):
            assert vals[i, 0] == 100 * a + 10 * b + c + 1000 * w
            assert vals[i, 1] == -(a + 1)
    # every row: two leading fields and four-space separated pairs with a trailing separator
    assert lines[3].endswith("    ")
    with pytest.raises(ValueError):
        R.write_response_tensor(0.0, omega, t[:1], 3, str(tmp_path), "bad.dat")
    with pytest.raises(ValueError):
        R.write_response_tensor(0.0, omega, t[:, :2], 3, str(tmp_path), "bad.dat")


def test_writer_nonfinite_spelling(tmp_path):
    t = np.array([[complex(np.nan, np.inf)]])
    path = R.write_response_tensor(0.0, [1.0], t, 1, str(tmp_path), "nf.dat")
    assert open(path).read().splitlines()[2].split()[2:] == ["NaN", "Inf"]


def test_energy_range_reference_cases():
    g = R.energy_range(0.0, 4.0, 0.02)
    assert len(g) == 201
    assert np.all(np.abs(g - np.array([float(Fraction(i, 50)) for i in range(201)])) <= 1e-15)
    assert np.all(np.abs(R.energy_range(0.0, 1.0, 0.3) - [0.0, 0.3, 0.6, 0.9]) <= 2e-16)
    assert list(R.energy_range(0.0, 0.3, 0.1)) == [0.0, 0.1, 0.2, 0.3]
    lin = R.energy_range(-0.2, 0.2, None, 41)
    assert np.all(np.abs(lin - np.array([float(Fraction(i, 100)) for i in range(-20, 21)])) <= 1e-16)
    assert list(R.energy_range(2.0, 2.0, 0.1)) == [2.0]
    assert list(R.energy_range(2.0, 2.0, None, 1)) == [2.0]


@pytest.mark.parametrize("args", [
    (0.0, 1.0, None, True), (0.0, 1.0, None, 1), (1.0, 1.0, None, 2), (0.0, 1.0, 0.0, None),
    (0.0, 1.0, -1.0, None), (0.0, float("inf"), 0.1, None), (False, 1.0, 0.1, None),
    (1e16, 1e16 + 4.0, 1.0, None), (0.0, 1.0, 1e-12, None), (0.0, 1.0, None, 10 ** 10),
    (2 ** 53, 2 ** 53 + 1, None, 1), (2 ** 53, 2 ** 53 + 1, None, 2),
])
def test_energy_range_rejections(args):
    with pytest.raises(ValueError):
        R.energy_range(*args)


def test_energy_range_count_exact_and_step_clamp():
    # count: each point the correctly rounded rational lerp, endpoints exact
    e = R.energy_range(0.1, 0.7, None, 7)
    assert e[0] == 0.1 and e[-1] == 0.7
    assert all(e[i] == float(Fraction(0.1) + (Fraction(0.7) - Fraction(0.1)) * i / 6) for i in range(7))
    # step: an overshooting last point within the tolerance is clamped to the maximum
    e = R.energy_range(0.0, 0.7, 0.1)
    assert len(e) == 8 and e[-1] == 0.7


def test_result_filenames():
    assert R.result_filename("GeS", "shift_current") == "GeS_sc_conv.dat"
    assert R.result_filename("GeS", "injection_current") == "GeS_ic_conv.dat"
    assert R.result_filename("GeS", "second_harmonic_generation") == "GeS_shg_conv.dat"
    with pytest.raises(ValueError):
        R.result_filename("GeS", "orbital_magnetization")
    with pytest.raises(ValueError):
        R.result_filename("GeS", "shift_current", method="Projector")
    with pytest.raises(ValueError):
        R.result_filename("GeS", "shift_current", calculation="kslice")


@pytest.mark.parametrize("output", ["total", "terms", "both"])
@pytest.mark.parametrize("response", ["both", "susceptibility", "conductivity"])
def test_second_harmonic_writer(tmp_path, output, response):
    rows, comps = 3, [(1, 1, 1), (3, 1, 2)]
    v = np.zeros((rows, len(comps), 7, 2), dtype=np.complex128)
    for r in range(rows):
        for c in range(len(comps)):
            for t in range(7):
                for q in range(2):
                    v[r, c, t, q] = complex(t + 1, -(q + 1) * (r + 1)) * (c + 1)
    coords = np.array([0.0, 0.5, 1.5])[:, None]
    paths = R.write_second_harmonic(str(tmp_path), coords, v, comps, output=output, response=response,
                                    domain_metadata=R.full_zone_metadata((4, 4, 4)))
    names = [os.path.basename(p) for p in paths]
    want = []
    if response != "conductivity":
        want += ["shg_chi_xxx.dat", "shg_chi_zxy.dat"]
    if response != "susceptibility":
        want += ["shg_sigma_xxx.dat", "shg_sigma_zxy.dat"]
    assert names == want
    for p in paths:
        q = 0 if "_chi_" in p else 1
        c = 0 if p.endswith("xxx.dat") else 1
        lines = open(p).read().splitlines()
        unit = "pm/V" if q == 0 else "A/V^2"
        assert lines[0] == f"# SHG {'chi' if q == 0 else 'sigma'} units={unit} component={'xxx' if c == 0 else 'zxy'}"
        assert lines[1] == "# full-BZ normalized susceptibility/conductivity"
        assert lines[2] == "# integration_bounds=nothing"
        assert lines[3] == "# integration_dimension=3"
        assert lines[6].startswith("# boundary_corrections=")
        labels = lines[7].split()[1:]
        assert labels[0] == "energy_eV"
        if output != "terms":
            assert labels[1:3] == ["total_re", "total_im"]
        if output != "total":
            assert labels[-14:-12] == ["interband_derivative_re", "interband_derivative_im"]
            assert labels[-2:] == ["semiclassical_re", "semiclassical_im"]
        for r, line in enumerate(lines[8:]):
            f = [float(x) for x in line.split()]
            assert f[0] == coords[r, 0]
            pos = 1
            if output != "terms":
                tot = v[r, c, :, q].sum()
                assert (f[1], f[2]) == (float("%.12e" % tot.real), float("%.12e" % tot.imag))
                pos = 3
            if output != "total":
                terms = np.array(f[pos:]).reshape(7, 2)
                assert np.array_equal(terms[:, 0], v[r, c, :, q].real)
                assert np.array_equal(terms[:, 1], v[r, c, :, q].imag)
        assert len(lines) == 8 + rows


def _write_cfg(tmp_path, body):
    shutil.copy(FIXTURE, tmp_path / "four_band_tb.dat")
    p = tmp_path / "cfg.toml"
    p.write_text(body)
    return str(p)


BASE = """
[model]
source = "wannier"
model_file = "four_band_tb.dat"

[response]
photon_energy_min_ev = 0.0
photon_energy_max_ev = 1.0
photon_energy_step_ev = 0.3
fermi_energy_ev = 0.1
temperature_k = 0.0

[response.sampling]
kind = "bz_mesh"
k_mesh = [3, 2, 2]
spatial_dimension = 3

[response.execution]
fourier_backend = "direct"

[response.tasks.sc]
kind = "SC"
method = "Conventional"

[response.tasks.ic]
kind = "injection_current"

[response.tasks.shg]
kind = "SHG"
[response.tasks.shg.numerics]
broadening = 0.03

[output]
output_root = "results/energy-step"
progress_enabled = false
"""


def test_load_run_config(tmp_path):
    cfg = R.load_run_config(_write_cfg(tmp_path, BASE))
    assert cfg.system_name == "four_band"
    assert cfg.model_file == str(tmp_path / "four_band_tb.dat")
    assert cfg.output_root == str(tmp_path / "results" / "energy-step")
    assert list(cfg.photon_energies) == list(R.energy_range(0.0, 1.0, 0.3))
    assert cfg.k_mesh == (3, 2, 2) and cfg.spatial_dimension == 3 and cfg.fermi_energy == 0.1
    assert [(t.task_id, t.quantity) for t in cfg.tasks] == [
        ("ic", "injection_current"), ("sc", "shift_current"), ("shg", "second_harmonic_generation")]
    assert cfg.tasks[2].numerics == {"broadening": 0.03}
    assert cfg.tasks[2].selectors == {} and cfg.tasks[0].selectors == {}
    body = BASE.replace('kind = "SHG"', 'kind = "SHG"\noutput = "both"\nresponse = "susceptibility"')
    shg = R.load_run_config(_write_cfg(tmp_path, body)).tasks[2]
    assert shg.selectors == {"output": "both", "response": "susceptibility"}
    # explicit list, explicit system name, default output root = the run file's directory
    body = BASE.replace("photon_energy_min_ev = 0.0\nphoton_energy_max_ev = 1.0\nphoton_energy_step_ev = 0.3",
                        "photon_energies_ev = [0.5, 1.5]")
    body = body.replace('output_root = "results/energy-step"', 'system_name = "GeS"')
    cfg = R.load_run_config(_write_cfg(tmp_path, body))
    assert list(cfg.photon_energies) == [0.5, 1.5] and cfg.system_name == "GeS"
    assert cfg.output_root == str(tmp_path)


@pytest.mark.parametrize("old,new", [
    ("spatial_dimension = 3", "spatial_dimension = 2"),
    ('kind = "bz_mesh"', 'kind = "kpath"'),
    ('kind = "SC"', 'kind = "orbital_magnetization"'),
    ('method = "Conventional"', 'method = "Projector"'),
    ("broadening = 0.03", "broadening = 0.03\nfinite_difference_step = 0.01"),
    ("progress_enabled = false", "progress_enabled = false\nunknown_key = 1"),
    ("photon_energy_step_ev = 0.3", "photon_energy_step_ev = 0.3\nphoton_energy_num_points = 4"),
    ('source = "wannier"', 'source = "openmx"'),
    ("temperature_k = 0.0", "temperature_k = -1.0"),
    ('kind = "SC"', 'kind = "SC"\noutput = "terms"'),
    ('kind = "SHG"', 'kind = "SHG"\noutput = "all"'),
    ("k_mesh = [3, 2, 2]", "k_mesh = [3, 0, 2]"),
    ("photon_energy_min_ev = 0.0\n", ""),
    ('fourier_backend = "direct"', 'fourier_backend = "fft"'),
])
def test_load_run_config_rejects(tmp_path, old, new):
    assert old in BASE
    with pytest.raises(ValueError):
        R.load_run_config(_write_cfg(tmp_path, BASE.replace(old, new, 1)))


@pytest.mark.parametrize("value,expected", [('"direct"', "direct"), ('"Mixed"', "mixed"), ('" AUTO "', "auto"),
                                            ('"-mixed-"', "mixed")])
def test_fourier_backend_names(tmp_path, value, expected):
    """fourier_backend is canonicalized like the reference's keys (Config.jl:437-442: lower case, blank and hyphen
    runs to one underscore, outer underscores stripped) and must then be direct, mixed or auto
    (FourierExecutionPlan.jl:63-67); without the key the backend is auto."""
    body = BASE.replace('fourier_backend = "direct"', f"fourier_backend = {value}")
    assert R.load_run_config(_write_cfg(tmp_path, body)).fourier_backend == expected
    body = BASE.replace('[response.execution]\nfourier_backend = "direct"\n', "")
    assert R.load_run_config(_write_cfg(tmp_path, body)).fourier_backend == "auto"


def test_cli_usage():
    assert R.main([]) == 2
    assert R.main(["check", "x.toml"]) == 2
