"""The RAPTOR command-line driver (raptor_amd/host/raptor): CLI, model.in,
dump readers and the output file formats of core.c:154-200 /
utilities.c:47-64."""
import os
import subprocess
import warnings

import numpy as np
import pytest

from conftest import M_UNIT, R_HIGH
from raptor_amd import build, harm, torus

DRIVER = build.DRIVER

MODEL_IN = """MBH         (g)      8.9505e39
M_UNIT      (g)      4.e20
ABSORPTION  (-)      1

IMG_WIDTH   (pixels) {W}
IMG_HEIGHT  (pixels) {H}
CAM_SIZE_X  (Rg)     40
CAM_SIZE_Y  (Rg)     40
FREQS_PER_DEC (-)    1
FREQ_MIN    (Hz)     2.3e11
FREQ_MAX    (Hz)     2.3e11
STEPSIZE    (-)      0.03
"""


@pytest.fixture(scope="module")
def driver():
    if not os.path.exists(DRIVER):
        pytest.fail("driver not built; run `python -m raptor_amd.build`")
    return DRIVER


def _run(driver, args, cwd):
    return subprocess.run([driver, *args], cwd=cwd, capture_output=True, text=True, timeout=300)


def test_usage_and_argument_errors(driver, tmp_path):
    r = _run(driver, [], tmp_path)
    assert r.returncode == 2 and "usage" in r.stderr
    (tmp_path / "model.in").write_text(MODEL_IN.format(W=8, H=8))
    r = _run(driver, ["model.in", "d", "1e23", "60", "10", "1", "0", "--bogus", "1"], tmp_path)
    assert r.returncode == 2 and "unknown flag" in r.stderr
    r = _run(driver, ["model.in", "d", "x", "60", "10", "1", "0"], tmp_path)
    assert r.returncode == 2 and "must be numbers" in r.stderr
    r = _run(driver, ["model.in", "d", "1e23", "60", "10", "1", "0", "--freqs", "1e11,abc"],
             tmp_path)
    assert r.returncode == 2 and "--freqs" in r.stderr
    r = _run(driver, ["model.in", "d", "1e23", "60", "10", "1", "0", "--integrator", "euler"],
             tmp_path)
    assert r.returncode == 2 and "--integrator" in r.stderr
    r = _run(driver, ["model.in", "d", "1e23", "60", "10", "1", "0", "--emission", "powerlaw"],
             tmp_path)
    assert r.returncode == 2 and "--emission" in r.stderr


def test_model_in_and_dump_errors(driver, tmp_path):
    (tmp_path / "bad.in").write_text("MBH (g) 8e39\n")
    r = _run(driver, ["bad.in", "d", "1e23", "60", "10", "1", "0"], tmp_path)
    assert r.returncode == 2 and "11 records" in r.stderr
    (tmp_path / "model.in").write_text(MODEL_IN.format(W=8, H=8))
    r = _run(driver, ["model.in", "missing", "1e23", "60", "10", "1", "0"], tmp_path)
    assert r.returncode == 1 and "can't open sim data file" in r.stderr
    (tmp_path / "short").write_text("0 4 4 0.3 0 0.1 0.25 0 0 0.9 1.4 0.9 0.5 0.5 1 10 0 0 0 "
                                    "0.001 0 0 1.3 40 0.3 0\n1 2 3\n")
    r = _run(driver, ["model.in", "short", "1e23", "60", "10", "1", "0"], tmp_path)
    assert r.returncode == 1 and "incomplete" in r.stderr


def test_driver_prints_accretion_fluxes(driver, tmp_path):
    """The line init_model prints after reading an ASCII dump
    (raptor_harm_model.c:171-182): dMact, the sum of gdet rho U^1 over radial
    zones 0..20, and Ladv, of gdet u U^1 U_0 over zones 20..39, each times
    dx3 dx2 (dx3 = 2 pi) and over 21; Macc = dMact M_unit / T_unit in
    Msun/year, T_unit = G MBH / c^3.  Printed before any GPU call, so checked
    here on the CPU, on a dump with random radial velocities (the torus
    itself has U^1 = 0)."""
    import re
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        d = torus.fishbone_moncrief(48, 24)
    rng = np.random.default_rng(3)
    d.extras["ucon"][1] = rng.uniform(-0.5, 0.5, d.extras["ucon"][1].shape)
    d.extras["ucov"][0] = rng.uniform(-1.2, -0.8, d.extras["ucov"][0].shape)
    harm.write_harm_ascii(str(tmp_path / "dump.txt"), d)
    (tmp_path / "model.in").write_text(MODEL_IN.format(W=4, H=4))
    r = _run(driver, ["model.in", "dump.txt", "%g" % M_UNIT, "60", "%g" % R_HIGH, "1", "0"],
             tmp_path)
    m = re.search(r"dMact: (\S+) Macc: (\S+) \[Msun/year\], Ladv: (\S+)\n", r.stderr)
    assert m, r.stderr
    assert "Done reading data" in r.stderr[m.end():]
    g, ex = d.grid, d.extras
    rho, uu = d.prims[0, :, :, 0], d.prims[1, :, :, 0]
    flux = ex["gdet"] * ex["ucon"][1]
    dM = (flux[:21] * rho[:21]).sum() * 2 * np.pi * g.dx2 / 21
    L = (flux[20:40] * uu[20:40] * ex["ucov"][0][20:40]).sum() * 2 * np.pi * g.dx2 / 21
    T_unit = 6.6742e-8 * 8.9505e39 / 2.99792458e10 ** 3
    macc = dM * M_UNIT / T_unit / (1.989e33 / 31536000.0)
    assert abs(dM) > 0 and abs(L) > 0
    assert float(m.group(1)) == pytest.approx(dM, rel=1e-5)
    assert float(m.group(2)) == pytest.approx(macc, rel=1e-6)
    assert float(m.group(3)) == pytest.approx(L, rel=1e-5)


def _expected_files(tmp_path, image, W, H, freqs, incl, dump_name, jf, t=0):
    """The reference's writers (core.c:163-198, utilities.c:51-61) applied to
    an image [W*H][nf]."""
    out = {}
    # sequential sum over pixels, as core.c:146-150 does
    spec = np.array([np.cumsum(image[:, f])[-1] for f in range(image.shape[1])])
    for f, nu in enumerate(freqs):
        name = "img_data_%d_%e_%.02f.dat" % (t, nu, incl)
        lines = ["%d %d %e %e %s %e %e %e %e\n" % (W, H, jf * spec[f], nu, dump_name, M_UNIT,
                                                   incl, 1.0, R_HIGH)]
        for i in range(W):
            for j in range(H):
                lines.append("%+.5e\n" % (jf * image[i + j * W, f]))
        out[name] = "".join(lines)
    out["spectrum_%d_%.02f.dat" % (t, incl)] = "".join(
        "%+.15e\t%+.15e\n" % (nu, jf * spec[f]) for f, nu in enumerate(freqs))
    return out


@pytest.mark.gpu
def test_driver_ascii_dump_end_to_end(driver, tmp_path):
    """ASCII HARM dump -> output files byte-identical to the reference's
    formats applied to the library's image."""
    from raptor_amd import rt
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        d = torus.fishbone_moncrief(48, 48)
    harm.write_harm_ascii(str(tmp_path / "dump.txt"), d)
    W, H = 20, 16
    (tmp_path / "model.in").write_text(MODEL_IN.format(W=W, H=H))
    r = _run(driver, ["model.in", "dump.txt", "%g" % M_UNIT, "60", "%g" % R_HIGH, "1", "0"],
             tmp_path)
    assert r.returncode == 0, r.stderr
    with rt.Tracer(0) as tr:
        tr.set_params(width=W, height=H, M_unit=M_UNIT, R_high=R_HIGH)
        tr.load_dump(d)
        img, _ = tr.render([60.0], [2.3e11])
        jf = tr.jansky_factor()
    want = _expected_files(tmp_path, img[0], W, H, [2.3e11], 60.0, "dump.txt", jf)
    for name, text in want.items():
        got = (tmp_path / "output" / name).read_text()
        assert got == text, name


@pytest.mark.gpu
def test_driver_binary_3d_dump_with_lists(driver, tmp_path):
    from raptor_amd import rt
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        d = torus.fishbone_moncrief(32, 24, 8)
    harm.write_prims_bin(str(tmp_path / "dump.bin"), d)
    W = H = 12
    (tmp_path / "model.in").write_text(MODEL_IN.format(W=W, H=H))
    freqs = [8.6e10, 2.3e11]
    r = _run(driver, ["model.in", "dump.bin", "%g" % M_UNIT, "60", "%g" % R_HIGH, "1", "0",
                      "--incl", "30,60", "--freqs", "8.6e10,2.3e11"], tmp_path)
    assert r.returncode == 0, r.stderr
    with rt.Tracer(0) as tr:
        tr.set_params(width=W, height=H, M_unit=M_UNIT, R_high=R_HIGH)
        tr.load_dump(d)
        img, _ = tr.render([30.0, 60.0], freqs)
        jf = tr.jansky_factor()
    for k, inc in enumerate((30.0, 60.0)):
        want = _expected_files(tmp_path, img[k], W, H, freqs, inc, "dump.bin", jf)
        for name, text in want.items():
            assert (tmp_path / "output" / name).read_text() == text, name


@pytest.mark.gpu
def test_driver_integrator_flag(driver, tmp_path):
    """--integrator rk4 traces with RK4 (int_method): the image equals the
    library's RK4 render and differs from the RK2 one."""
    from raptor_amd import rt
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        d = torus.fishbone_moncrief(48, 48)
    harm.write_harm_ascii(str(tmp_path / "dump.txt"), d)
    W, H = 16, 16
    (tmp_path / "model.in").write_text(MODEL_IN.format(W=W, H=H))
    r = _run(driver, ["model.in", "dump.txt", "%g" % M_UNIT, "60", "%g" % R_HIGH, "1", "0",
                      "--integrator", "rk4"], tmp_path)
    assert r.returncode == 0, r.stderr
    with rt.Tracer(0) as tr:
        tr.set_params(width=W, height=H, M_unit=M_UNIT, R_high=R_HIGH, integrator=rt.INT_RK4)
        tr.load_dump(d)
        img4, _ = tr.render([60.0], [2.3e11])
        tr.set_params(width=W, height=H, M_unit=M_UNIT, R_high=R_HIGH)
        img2, _ = tr.render([60.0], [2.3e11])
        jf = tr.jansky_factor()
    want = _expected_files(tmp_path, img4[0], W, H, [2.3e11], 60.0, "dump.txt", jf)
    for name, text in want.items():
        assert (tmp_path / "output" / name).read_text() == text, name
    assert not np.array_equal(img4, img2)


@pytest.mark.gpu
def test_driver_emission_flag(driver, tmp_path):
    """--emission kappa renders with the kappa fits: the image equals the
    library's kappa render and differs from the thermal one."""
    from raptor_amd import rt
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        d = torus.fishbone_moncrief(48, 48)
    harm.write_harm_ascii(str(tmp_path / "dump.txt"), d)
    W, H = 16, 16
    (tmp_path / "model.in").write_text(MODEL_IN.format(W=W, H=H))
    r = _run(driver, ["model.in", "dump.txt", "%g" % M_UNIT, "60", "%g" % R_HIGH, "1", "0",
                      "--emission", "kappa"], tmp_path)
    assert r.returncode == 0, r.stderr
    with rt.Tracer(0) as tr:
        tr.set_params(width=W, height=H, M_unit=M_UNIT, R_high=R_HIGH, emission=rt.EMIS_KAPPA)
        tr.load_dump(d)
        imgk, _ = tr.render([60.0], [2.3e11])
        tr.set_params(width=W, height=H, M_unit=M_UNIT, R_high=R_HIGH)
        imgt, _ = tr.render([60.0], [2.3e11])
        jf = tr.jansky_factor()
    want = _expected_files(tmp_path, imgk[0], W, H, [2.3e11], 60.0, "dump.txt", jf)
    for name, text in want.items():
        assert (tmp_path / "output" / name).read_text() == text, name
    assert not np.array_equal(imgk, imgt)
