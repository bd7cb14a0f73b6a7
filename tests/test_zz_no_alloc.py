"""m1_gamma_pass_count_noalloc against m1_gamma_pass_count (GPU).

The no-allocation entry point takes device-resident dose, reference and
point arrays (here allocated by the test with hipMalloc from the HIP runtime
the library itself links, so that only one runtime lives in the process) and
stages the search offsets and counters in the library's fixed device
scratch.  Its pass count must equal the allocating entry point's on the same
inputs, call after call (the scratch keeps the last call's contents, so the
second, smaller call checks that the counters are cleared and that the
smaller offset table is what is read), and its limits must be errors.
"""

import ctypes as C

import numpy as np
import pytest

from proton_solver import _native, metrics

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class DeviceArrays:
    """Caller-owned device buffers (hipMalloc of the library's HIP runtime)."""

    def __init__(self):
        _native.load_library()
        self.hip = C.CDLL("libamdhip64.so.7")  # the runtime libm1march.so is linked to
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.hip.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.hip.hipStreamDestroy.argtypes = [C.c_void_p]
        n = C.c_int(0)
        if self.hip.hipGetDeviceCount(C.byref(n)) != 0 or n.value < 1:
            pytest.skip("no HIP device")
        assert self.hip.hipSetDevice(0) == 0
        self.ptrs = []

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), a.nbytes) == 0
        self.ptrs.append(p)
        assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        return p.value

    def stream(self):
        s = C.c_void_p()
        assert self.hip.hipStreamCreate(C.byref(s)) == 0
        self.ptrs.append(("stream", s))
        return s.value

    def close(self):
        for p in self.ptrs:
            if isinstance(p, tuple):
                self.hip.hipStreamDestroy(p[1])
            else:
                self.hip.hipFree(p)
        self.ptrs = []


@pytest.fixture
def dev():
    d = DeviceArrays()
    yield d
    d.close()


def _case(shape, seed, shift):
    rng = np.random.default_rng(seed)
    x, y, z = np.meshgrid(*[np.arange(n, dtype=float) for n in shape], indexing="ij")
    c = [n / 2 for n in shape]
    ref = np.exp(-((x - c[0]) ** 2 + (y - c[1]) ** 2) / 18.0) * (1 + 0.02 * z)
    dose = np.exp(-((x - c[0] - shift) ** 2 + (y - c[1]) ** 2) / 18.0) * (1 + 0.02 * z)
    dose *= 1 + 0.01 * rng.standard_normal(shape)
    return dose, ref


def _device_count(dev, dose, ref, spacing=(0.001, 0.001, 0.001), interp_fraction=10,
                  stream=0):
    dose_s, ref_s, pts, scaled, dist_terms, dd = metrics._gamma_setup(
        dose, ref, spacing, 1.0, 1.0, 10.0, interp_fraction, 1.1)
    expect = _native.gamma_pass_count(0, dose_s, ref_s, pts, scaled, dist_terms, dd)
    got = _native.gamma_pass_count_noalloc(
        0, dev.upload(dose_s.astype(np.float64)), dev.upload(ref_s.astype(np.float64)),
        dose_s.shape, dev.upload(pts.astype(np.int32)), len(pts), scaled, dist_terms, dd, stream)
    return got, expect, len(pts), len(scaled)


def test_noalloc_gamma_equals_allocating_entry_twice(dev):
    dose, ref = _case((40, 36, 48), 0, 0.6)
    got, expect, n_pts, n_offs = _device_count(dev, dose, ref, stream=dev.stream())
    print(f"noalloc gamma call 1: {n_pts} points, {n_offs} offsets, {got} / {expect}", flush=True)
    assert 0 < expect < n_pts and got == expect
    # a smaller grid, fewer points and a smaller offset table (a coarser
    # search step) right after the larger call, on the default stream
    dose2, ref2 = _case((22, 20, 26), 1, 1.3)
    got2, expect2, n_pts2, n_offs2 = _device_count(dev, dose2, ref2,
                                                   spacing=(0.002, 0.002, 0.002),
                                                   interp_fraction=5)
    print(f"noalloc gamma call 2: {n_pts2} points, {n_offs2} offsets, {got2} / {expect2}",
          flush=True)
    assert n_pts2 < n_pts and n_offs2 < n_offs
    assert 0 < expect2 < n_pts2 and got2 == expect2


def test_noalloc_gamma_limits_are_errors(dev):
    dose, ref = _case((12, 12, 12), 2, 0.3)
    d = dev.upload(dose)
    r = dev.upload(ref)
    pts = dev.upload(np.array([[5, 5, 5], [6, 6, 6]], dtype=np.int32))
    n_max = _native.GAMMA_NOALLOC_MAX_OFFSETS
    offs = np.zeros((n_max + 1, 3))
    dist = np.zeros(n_max + 1)
    with pytest.raises(ValueError, match="exceeds"):
        _native.gamma_pass_count_noalloc(0, d, r, dose.shape, pts, 2, offs, dist, 0.01)
    # a point outside the grid is skipped by the kernel and reported
    bad = dev.upload(np.array([[5, 5, 5], [12, 0, 0]], dtype=np.int32))
    with pytest.raises(ValueError, match="outside the grid"):
        _native.gamma_pass_count_noalloc(0, d, r, dose.shape, bad, 2, offs[:1], dist[:1], 0.01)
    # the next call starts from cleared counters
    n = _native.gamma_pass_count_noalloc(0, d, r, dose.shape, pts, 2, offs[:1], dist[:1], 1e9)
    assert n == 2
