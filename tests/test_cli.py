"""The C++ front end ``tt_demo`` (options and outputs of c++/src/demo.cpp,
file naming of c++/src/transform.cpp:61-72,107-111, netpbm input of
c++/src/auxiliary.cpp:26-99).

CPU tests cover option parsing and input validation, which end before any
device call.  GPU tests run the transform through the C++ drop-in header and
compare the CSV files it writes with the Python path's results printed the
same way (6 significant digits), and with the oracle within the CSV's
rounding."""
import os
import subprocess

import numpy as np
import pytest

from tracetransform_amd import build as tt_build
from tracetransform_amd import inputs

DEMO = tt_build.DEMO


def run_demo(args, cwd):
    if not os.path.exists(DEMO):
        pytest.fail("tt_demo is not built (run __graft_entry__.build())")
    return subprocess.run([DEMO, *args], cwd=cwd, capture_output=True, text=True, timeout=120)


def write_pgm(path, gray, comment=True):
    rows, cols = gray.shape
    with open(path, "w") as f:
        f.write("P2\n")
        if comment:
            f.write("# written by the test\n")
        f.write(f"{cols} {rows}\n255\n")
        for r in range(rows):
            f.write(" ".join(str(int(v)) for v in gray[r]) + "\n")


def write_ppm(path, rgb):
    rows, cols, _ = rgb.shape
    with open(path, "w") as f:
        f.write(f"P3\n{cols} {rows}\n255\n")
        for r in range(rows):
            f.write(" ".join(str(int(v)) for v in rgb[r].reshape(-1)) + "\n")


def read_csv(path):
    with open(path) as f:
        return [[tok.strip() for tok in line.rstrip("\n").split(",")] for line in f]


def fmt6(a):
    """Rows of a 2-D array printed as a default std::ostream prints floats."""
    return [[f"{float(v):.6g}" for v in row] for row in np.asarray(a)]


def gray_image(rows, cols, seed):
    img = inputs.blob_image(cols, seed, rows=rows)
    return np.clip(np.round(img / img.max() * 255.0), 0, 255).astype(np.int64)


# ---- CPU: option parsing and validation (no device call is reached) --------

def test_help(tmp_path):
    r = run_demo(["--help"], tmp_path)
    assert r.returncode == 0 and "--t-functional" in r.stdout


@pytest.mark.parametrize("args,msg", [
    (["-m", "calculate", "x.pgm"], "'--t-functional' is required"),
    (["-T", "0", "x.pgm"], "'--mode' is required"),
    (["-T", "0", "-m", "calculate"], "'--inputs' is required"),
    (["-T", "8", "-m", "calculate", "x.pgm"], "--t-functional' is invalid"),
    (["-T", "0", "-P", "H", "-m", "calculate", "x.pgm"], "--p-functional' is invalid"),
    (["-T", "0", "-P", "Hx", "-m", "calculate", "x.pgm"], "--p-functional' is invalid"),
    (["-T", "0", "-m", "compute", "x.pgm"], "--mode' is invalid"),
    (["-T", "0", "-a", "x", "-m", "calculate", "x.pgm"], "--angle' is invalid"),
    (["-T", "0", "--colour", "-m", "calculate", "x.pgm"], "unrecognised option"),
])
def test_usage_errors(tmp_path, args, msg):
    r = run_demo(args, tmp_path)
    assert r.returncode == 1
    assert msg in r.stderr


def test_input_errors(tmp_path):
    r = run_demo(["-T", "0", "-P", "1", "-P", "H2", "-m", "calculate", "x.pgm"], tmp_path)
    assert r.returncode == 1 and "Cannot mix" in r.stderr
    r = run_demo(["-T", "0", "-m", "calculate", "missing.pgm"], tmp_path)
    assert r.returncode == 1 and "does not exist" in r.stderr
    (tmp_path / "a.png").write_text("x")
    r = run_demo(["-T", "0", "-m", "calculate", "a.png"], tmp_path)
    assert r.returncode == 1 and "Unrecognized input file format" in r.stderr
    (tmp_path / "b.pgm").write_text("P5\n2 2\n255\n")
    r = run_demo(["-T", "0", "-m", "calculate", "b.pgm"], tmp_path)
    assert r.returncode == 1 and "Invalid Netpbm magic" in r.stderr
    (tmp_path / "c.pgm").write_text("P2\n2 2\n255\n1 2 3\n")
    r = run_demo(["-T", "0", "-m", "calculate", "c.pgm"], tmp_path)
    assert r.returncode == 1 and "Error processing file" in r.stderr
    write_pgm(tmp_path / "d.pgm", np.ones((3, 3), dtype=np.int64))
    r = run_demo(["-T", "0", "-m", "benchmark", "d.pgm"], tmp_path)
    assert r.returncode == 1 and "'--iterations' is required" in r.stderr


# ---- GPU: outputs of the C++ drop-in path ---------------------------------

def python_signatures(gray, step, tcodes, pspecs, orthonormal=False):
    from tracetransform_amd import PFunctionalWrapper, TFunctionalWrapper, transform
    import torch
    img = torch.from_numpy((gray / 255.0).astype(np.float32)).cuda()
    tw = [TFunctionalWrapper.parse(str(t)) for t in tcodes]
    pw = [PFunctionalWrapper.parse(p) for p in pspecs]
    tr = transform.Transformer(img, "", step, orthonormal)
    sig = tr.getTransform(tw, pw)
    n = tr.image.shape[0]
    sino = transform.getSinograms(tr.image, step, tw)
    return (sig.cpu().numpy() if sig is not None else None,
            [s.cpu().numpy() for s in sino], n, pw)


@pytest.mark.gpu
def test_calculate_writes_reference_files(tmp_path, gpu, oracle):
    gray = gray_image(37, 52, seed=11)
    write_pgm(tmp_path / "blob.pgm", gray)
    step, tcodes, pspecs = 4, [0, 3, 6], ["1", "2", "3"]
    r = run_demo(["-m", "calculate", "-v", "-a", str(step)]
                 + sum((["-T", str(t)] for t in tcodes), [])
                 + sum((["-P", p] for p in pspecs), []) + ["blob.pgm"], tmp_path)
    assert r.returncode == 0, r.stderr
    sig, sinos, n, _ = python_signatures(gray, step, tcodes, pspecs)
    A = 360 // step
    names = ["Radon", "T3", "T6"]
    for t, name in enumerate(names):
        got = read_csv(tmp_path / f"blob-{name}.csv")
        assert len(got) == n and all(len(row) == A for row in got)
        assert got == fmt6(sinos[t]), name
        assert (tmp_path / f"blob-{name}.pgm").read_text().startswith(f"P2\n{A} {n}\n255\n")
    got = read_csv(tmp_path / "blob.csv")
    assert got == fmt6(sig)
    # against the oracle's getTransform on the same padded image; P2 is an
    # exact pick and the others agree to the CSV's 6 digits up to FP32 noise
    # amplified by the z-score
    img_cm = (gray / 255.0).astype(np.float32).T.reshape(-1)
    padded = oracle.pad(img_cm, gray.shape[0], gray.shape[1])
    want = oracle.transform(padded, n, A, step, tcodes, [(int(p),) for p in pspecs])
    g = np.array([[float(v) for v in row] for row in got], dtype=np.float64)
    assert np.max(np.abs(g - want.T)) <= 1e-3


@pytest.mark.gpu
def test_quiet_calculate_writes_only_signatures(tmp_path, gpu):
    gray = gray_image(20, 20, seed=12)
    write_pgm(tmp_path / "q.pgm", gray, comment=False)
    r = run_demo(["-m", "calculate", "-T", "1", "-P", "2", "-a", "10", "q.pgm"], tmp_path)
    assert r.returncode == 0, r.stderr
    assert sorted(os.listdir(tmp_path)) == ["q.csv", "q.pgm"]


@pytest.mark.gpu
def test_rgb_components(tmp_path, gpu):
    rgb = np.stack([gray_image(24, 30, seed=s) for s in (21, 22, 23)], axis=-1)
    write_ppm(tmp_path / "c.ppm", rgb)
    r = run_demo(["-m", "calculate", "-T", "0", "-T", "5", "-P", "1", "-P", "3", "-a", "6",
                  "c.ppm"], tmp_path)
    assert r.returncode == 0, r.stderr
    for k in range(3):
        sig, _, _, _ = python_signatures(rgb[:, :, k], 6, [0, 5], ["1", "3"])
        assert read_csv(tmp_path / f"c_c{k + 1}.csv") == fmt6(sig)


@pytest.mark.gpu
def test_orthonormal_hermite(tmp_path, gpu):
    gray = gray_image(30, 30, seed=31)
    write_pgm(tmp_path / "h.pgm", gray)
    r = run_demo(["-m", "calculate", "-T", "0", "-T", "4", "-P", "H1", "-P", "H3", "-a", "10",
                  "h.pgm"], tmp_path)
    assert r.returncode == 0, r.stderr
    sig, _, _, pw = python_signatures(gray, 10, [0, 4], ["H1", "H3"], orthonormal=True)
    assert all(p.arguments.center is not None for p in pw)
    assert read_csv(tmp_path / "h.csv") == fmt6(sig)


@pytest.mark.gpu
def test_benchmark_and_profile_modes(tmp_path, gpu):
    gray = gray_image(48, 48, seed=41)
    write_pgm(tmp_path / "b.pgm", gray)
    r = run_demo(["-m", "benchmark", "-n", "3", "-T", "0", "-T", "1", "-P", "1", "-a", "5",
                  "b.pgm"], tmp_path)
    assert r.returncode == 0, r.stderr
    lines = [ln for ln in r.stdout.splitlines() if ln.startswith("t_")]
    assert [ln.split("=")[0] for ln in lines] == ["t_1", "t_2", "t_3"]
    assert all(float(ln.split("=")[1]) > 0 for ln in lines)
    r = run_demo(["-m", "profile", "-T", "0", "-P", "1", "b.pgm"], tmp_path)
    assert r.returncode == 0, r.stderr
    assert sorted(os.listdir(tmp_path)) == ["b.pgm"]
