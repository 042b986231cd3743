"""The spmv_all / spmv_cocktail drivers: usage text on the CPU, and on the
GPU every spmv_all method plus spmv_cocktail on small matrices, with every
variant's one-launch check (host FP64 product, condition-aware relative
error <= 1e-5) required to pass."""
import os
import re
import subprocess

import pytest

from clspmv_amd import matrix as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "clspmv_amd", "bin")


def _run(args, timeout=120):
    return subprocess.run(args, capture_output=True, text=True, timeout=timeout)


@pytest.mark.parametrize("exe", ["spmv_all", "spmv_cocktail"])
def test_usage(exe):
    r = _run([os.path.join(BIN, exe)])
    assert r.returncode == 0 and r.stdout.startswith("Usage:"), r.stdout + r.stderr


@pytest.fixture(scope="module")
def mtx(tmp_path_factory):
    d = tmp_path_factory.mktemp("mtx")
    out = {}
    for name, coo in (("lap256", M.gen_laplacian2d(256, 256)),
                      ("blocks4", M.gen_block_rows(3000, 4, 6, 200, 3)),
                      ("cocktail", M.gen_cocktail_composite(65536, 6))):
        path = str(d / f"{name}.mtx")
        coo.write_mmf(path)
        out[name] = path
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("method", list(range(12)))
def test_spmv_all_every_method(mtx, method):
    path = mtx["blocks4"] if method in (7, 8, 10, 11) else mtx["lap256"]
    print(f"spmv_all {path} {method}", flush=True)
    r = _run([os.path.join(BIN, "spmv_all"), path, str(method), "5"])
    assert r.returncode == 0, r.stdout + r.stderr
    if method == 0:
        assert "GB/s" in r.stdout and "per launch" in r.stdout
        return
    errs = [float(e) for e in re.findall(r"max_rel_err ([0-9.e+-]+)", r.stdout)]
    assert errs, r.stdout
    assert "CHECK FAILED" not in r.stdout and max(errs) <= 1e-5, r.stdout
    assert "best:" in r.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["lap256", "cocktail"])
def test_spmv_cocktail(mtx, name):
    r = _run([os.path.join(BIN, "spmv_cocktail"), mtx[name], "20"])
    assert r.returncode == 0, r.stdout + r.stderr
    assert re.search(r"check: max_rel_err [0-9.e+-]+\n", r.stdout) and "Cocktail time" in r.stdout, r.stdout
