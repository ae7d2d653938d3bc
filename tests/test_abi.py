"""CPU tests of the product library's boundary (no GPU, no compute calls):
libscft_hip.so loads, exports every entry point include/scft.h declares, and
its host-only logic (contour counts, parameter validation, defaults) behaves;
the seeded-field generator reproduces the published SplitMix64 sequence."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "scft.h")


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(scft_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_the_abi(pkg):
    names = header_functions()
    assert set(names) == set(pkg.ABI_SYMBOLS), names


def test_library_exports_every_symbol(pkg):
    L = ctypes.CDLL(pkg.LIB_PATH)
    for name in header_functions():
        assert hasattr(L, name), name


def test_contour_counts_match_reference(pkg):
    assert pkg.contour_counts(0.5, 0.01) == (50, 50)
    assert pkg.contour_counts(0.3, 0.005) == (60, 140)
    assert pkg.contour_counts(0.35, 0.005) == (70, 130)
    assert pkg.contour_counts(0.375, 0.01) == (37, 62)
    with pytest.raises(pkg.ScftError):
        pkg.contour_counts(0.5, 0.0)


def test_default_params(pkg):
    p = pkg.default_params()
    assert p.wopt == 0.05 and p.wcmp == 0.10  # cuda_aid.cu:395-396
    assert p.pi == 3.1415926535                # struct.h:11
    assert p.precision == pkg.FP64 and p.mixing == pkg.MIX_SIMPLE and p.batch == 1
    assert p.use_graph == 1


@pytest.mark.parametrize("kw,code", [
    (dict(nx=0), 1), (dict(lx=-1.0), 1), (dict(ds=0.0), 1), (dict(fA=1.0), 1), (dict(fA=0.0), 1),
    (dict(batch=0), 1), (dict(precision=7), 1), (dict(mixing=3), 1),
    (dict(mixing=1, anderson_m=33), 1), (dict(fA=0.005, ds=0.01), 1),
    (dict(nx=7), 2), (dict(nx=14), 2), (dict(nx=1, ny=1, nz=1), 1),
])
def test_create_rejects_bad_params(pkg, kw, code):
    """Validation happens before any device call, so it is testable without a GPU."""
    p = pkg.default_params(**kw)
    with pytest.raises(pkg.ScftError) as e:
        pkg.Solver(p)
    assert e.value.code == code, str(e.value)


def test_splitmix64_published_sequence(pkg):
    # SplitMix64 with seed 1234567 (the reference implementation's test vector)
    i = np.arange(1, 6, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(1234567) + i * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    assert [int(v) for v in z] == [6457827717110365317, 3203168211198807973, 9817491932198370423,
                                   4593380528125082431, 16408922859458223821]
    u = pkg.fields.splitmix_uniform(1234567, 5)
    assert np.array_equal(u, (z >> np.uint64(11)).astype(np.float64) / 2.0 ** 53)


def test_seeded_fields_shapes(pkg):
    wA, wB = pkg.fields.seeded_fields(8, 6, 4, 2.0, 3.0, 4.0, "gyroid", 1.0, 0.1, 3)
    assert wA.shape == (192,) and wB.shape == (192,)
    S = pkg.fields.seed_shape("gyroid", 8, 6, 4, 2.0, 3.0, 4.0)
    # w_A + w_B is pure noise, w_A - w_B = 2 S + noise
    assert np.max(np.abs(wA + wB)) <= 0.2 + 1e-12
    assert np.max(np.abs((wA - wB) - 2 * S)) <= 0.2 + 1e-12
    w2 = pkg.fields.seeded_fields(8, 6, 4, 2.0, 3.0, 4.0, "gyroid", 1.0, 0.1, 3)
    assert np.array_equal(wA, w2[0])


def test_propagate_noalloc_rejects_sizes_beyond_its_scratch(pkg):
    """the no-allocation propagate refuses batch * N above its static scratch
    (checked before any device call)"""
    L = pkg.lib()
    p = pkg.default_params(nx=64, ny=64, nz=64, batch=2)
    dummy = ctypes.c_void_p(16)
    rc = L.scft_propagate_noalloc(ctypes.byref(p), 0, 0, 1, dummy, None, dummy, None)
    assert rc == pkg.SCFT_E_INVALID
    assert b"scratch limit" in L.scft_last_error(None)
