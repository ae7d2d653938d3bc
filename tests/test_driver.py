"""CPU tests of the C++ driver (no GPU): ab.txt parsing with the reference's
sample file, the seeded phi_<n>.dat writer (-gen) against the Python
generator, the field-file reader/writer round trip, the -kernal flag handling
and the VTK converter."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN_DIR, PKG_DIR

DRIVER = os.path.join(PKG_DIR, "scft")
VTK = os.path.join(PKG_DIR, "scft_vtk")

pytestmark = pytest.mark.skipif(not (os.path.exists(DRIVER) and os.path.exists(VTK)),
                                reason="driver not built (python scft_ver1.01_amd/build.py)")


def run(args, cwd):
    env = dict(os.environ, LD_LIBRARY_PATH=PKG_DIR)
    return subprocess.run([DRIVER] + args, cwd=cwd, env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                          text=True, timeout=120)


def write_ab(path, intag=0, n=(8, 6, 4), extra=""):
    with open(path, "w") as f:
        f.write(f"intag={intag}\nxN=15.0\nfa=0.500\nlx=4.0, ly=3.0,  lz=2.0\n"
                f"Nx={n[0]}, Ny={n[1]}, Nz={n[2]}\nds0=0.01\nbatch=2\n{extra}")


def test_reference_ab_txt_parses(tmp_path):
    """The reference's own sample configuration (tests/golden/ab_reference.txt,
    a copy of its ab.txt data): parsed, then -gen writes the seeded inputs."""
    shutil.copy(os.path.join(GOLDEN_DIR, "ab_reference.txt"), tmp_path / "ab.txt")
    # intag=1024 wants phi_1.dat, which the reference does not ship: reported, exit 1
    r = run(["-kernal=GPU", "-step=10"], str(tmp_path))
    assert r.returncode == 1 and "phi_1.dat did not exist" in r.stdout


def test_gen_matches_python_generator(pkg, tmp_path):
    # the solver keys (autotune, checkpoint, anderson_reg) parse alongside the generator's
    write_ab(tmp_path / "ab.txt", extra="field=gyroid\namp=1.0\nnoise=0.1\nseed=9\nautotune=0\ncheckpoint=1\n"
                                      "anderson_reg=0.01\nmixing=anderson\n")
    r = run(["-gen"], str(tmp_path))
    assert r.returncode == 0, r.stdout
    for i in range(2):
        rows = np.loadtxt(tmp_path / f"phi_{i + 1}.dat", skiprows=2)
        wA, wB = pkg.fields.seeded_fields(8, 6, 4, 4.0, 3.0, 2.0, "gyroid", 1.0, 0.1, 9 + i)
        assert rows.shape == (192, 4)
        assert np.max(np.abs(rows[:, 2] - wA)) < 1e-14 and np.max(np.abs(rows[:, 3] - wB)) < 1e-14
        head = open(tmp_path / f"phi_{i + 1}.dat").readline()
        assert head.strip() == "Nx=8, Ny=6, Nz=4"


def test_flags(tmp_path):
    write_ab(tmp_path / "ab.txt")
    r = run(["-kernal=CPU"], str(tmp_path))
    assert r.returncode == 2 and "In this programm, we use CPU." in r.stdout
    r = run(["-kernal=TPU"], str(tmp_path))
    assert r.returncode == 2
    r = run(["-gen"], str(tmp_path))
    assert "defalt set as 100" in r.stdout  # no -step: the reference's message
    r = run(["-GPU_N=2", "-step=5"], str(tmp_path))
    assert r.returncode == 2


def test_missing_and_malformed_config(tmp_path):
    r = run(["-step=3"], str(tmp_path))
    assert r.returncode == 1 and "ab.txt did not exist" in r.stdout
    (tmp_path / "ab.txt").write_text("intag=0\nxN=15\n")
    r = run(["-step=3"], str(tmp_path))
    assert r.returncode == 1 and "malformed" in r.stdout
    write_ab(tmp_path / "ab.txt", extra="bogus=1\n")
    r = run(["-step=3"], str(tmp_path))
    assert r.returncode == 1 and "unknown key" in r.stdout


def test_vtk_converter(tmp_path):
    n = (4, 3, 2)
    rows = np.arange(24 * 4, dtype=float).reshape(24, 4) / 100.0
    with open(tmp_path / "pha.dat", "w") as f:
        f.write("Nx=4, Ny=3, Nz=2\ndx=0.5, dy=0.25, dz=1\n")
        for r_ in rows:
            f.write("%.17g %.17g %.17g %.17g\n" % tuple(r_))
    env = dict(os.environ, LD_LIBRARY_PATH=PKG_DIR)
    r = subprocess.run([VTK, "pha.dat", "o.vtk"], cwd=tmp_path, env=env, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=60)
    assert r.returncode == 0, r.stdout
    text = open(tmp_path / "o.vtk").read().split("\n")
    assert text[0] == "# vtk DataFile Version 3.0"
    assert "DIMENSIONS 4 3 2" in text and "SPACING 0.500000 0.250000 1.000000" in text
    assert "POINT_DATA 24" in text
    vals = np.array(" ".join(text[text.index("LOOKUP_TABLE default") + 2:]).split(), dtype=float)
    assert np.allclose(vals, rows[:, 0] - rows[:, 1], atol=5e-5)


@pytest.mark.gpu
@pytest.mark.parametrize("precision,keys", [("fp64", ""), ("fp32x", ""),
                                            ("fp64", "checkpoint=1\nmixing=anderson\nanderson_m=4\n")],
                         ids=["fp64", "fp32x", "fp64-checkpoint-anderson"])
def test_driver_run_matches_library(pkg, tmp_path, precision, keys):
    """Full driver run on the GPU: F line every 10 iterations (the reference's
    printf format), pha files after 100, time= line; values equal the library's
    (the FP64 form; through precision=fp32x the FP32X form; through checkpoint=1
    and mixing=anderson the checkpointed store and Anderson mixing)."""
    write_ab(tmp_path / "ab.txt", n=(16, 16, 16), extra="field=lamellar\namp=1.0\nnoise=0.1\nseed=3\nexact_output=1\n"
             # FP32X: the library's default tiling, which the comparison below uses (another tiling
             # rounds differently at ~1e-7); FP64 keeps the driver's autotuning (rounding at ~1e-14)
             f"precision={precision}\n" + ("autotune=0\n" if precision == "fp32x" else "") + keys)
    r = run(["-kernal=GPU", "-step=20"], str(tmp_path))
    assert r.returncode == 0, r.stdout
    lines = [l for l in r.stdout.splitlines() if l.startswith("GPU 0 batch")]
    assert len(lines) == 4  # 2 instances x iterations 10, 20
    assert lines[0].startswith("GPU 0 batch 0:    10 : ")
    assert any(l.startswith("time=") for l in r.stdout.splitlines())
    vals = [float(v) for v in lines[2].split(":")[2].replace(",", " ").split()]
    kw = dict(checkpoint=1, mixing=pkg.MIX_ANDERSON, anderson_m=4) if keys else {}
    p = pkg.default_params(nx=16, ny=16, nz=16, lx=4.0, ly=3.0, lz=2.0, chiN=15.0, fA=0.5, ds=0.01,
                           precision={"fp64": pkg.FP64, "fp32x": pkg.FP32X}[precision], **kw)
    wA, wB = pkg.fields.seeded_fields(16, 16, 16, 4.0, 3.0, 2.0, "lamellar", 1.0, 0.1, 3)
    with pkg.Solver(p) as s:
        s.set_fields(wA, wB)
        _, fe = s.iterate(20)
    assert np.allclose(vals, fe[0], rtol=5e-8, atol=1e-12)  # %.8e printing
    assert not os.path.exists(tmp_path / "pha_1.dat")  # written every 100 iterations


def test_precision_key(tmp_path):
    """precision= takes fp64/double, fp32/float and fp32x; anything else is an
    input error naming the key, not a silent FP64 run."""
    for v in ("fp64", "double", "fp32", "float", "fp32x"):
        write_ab(tmp_path / "ab.txt", extra=f"precision={v}\n")
        r = run(["-gen"], str(tmp_path))
        assert r.returncode == 0, (v, r.stdout)
    write_ab(tmp_path / "ab.txt", extra="precision=half\n")
    r = run(["-gen"], str(tmp_path))
    assert r.returncode != 0 and "precision must be fp64, fp32 or fp32x" in r.stdout, r.stdout
