"""The FP64 kernels divide with the hardware reciprocal, two Newton steps and
one correction, and take the sound speed's square root with the compiler's
IEEE sequence minus its scaling of tiny arguments (wave-multiphase_amd/csrc/
wamp_device.h: ddiv, sq).  Bitwise FP64 parity with the oracle rests on both
being exactly the IEEE results over their documented operand ranges; these
tests run the device sequences through the wamp_selftest_arith hook beside
the compiler's a / b and sqrt and compare bit patterns, with numpy's host
results as a third opinion."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# divisors the solver uses: densities, c^2, mixture gamma - 1, tanh(0.9) in THINC
SPECIALS = [1.0e-10, 2.0e-10, 1.0e-20, 0.4, 3.4, 1000.0, 0.7162978701990245,
            6000.0 * 4.4, 1.0, 2.0, 3.0, 1.0 / 3.0]


def _doubles(rng, n, emin, emax):
    mant = rng.integers(0, 1 << 52, n, dtype=np.uint64)
    exp = rng.integers(emin + 1023, emax + 1024, n).astype(np.uint64)
    sign = rng.integers(0, 2, n).astype(np.uint64)
    return ((sign << np.uint64(63)) | (exp << np.uint64(52)) | mant).view(np.float64)


def _all_ones(rng, x):
    """The same values with (nearly) all-ones mantissas, the hardest divisors."""
    bits = x.view(np.uint64) | np.uint64((1 << 52) - 1)
    bits ^= rng.integers(0, 8, x.size).astype(np.uint64)
    return bits.view(np.float64)


def _run(a, b):
    import torch

    from wamp import _lib

    lib = _lib.load()
    n = a.size
    da = torch.from_numpy(a).cuda()
    db = torch.from_numpy(b).cuda()
    out = torch.empty(4 * n, dtype=torch.float64, device="cuda")
    st = torch.cuda.current_stream().cuda_stream
    _lib.check(lib.wamp_selftest_arith(da.data_ptr(), db.data_ptr(), out.data_ptr(), n, st),
               "wamp_selftest_arith")
    torch.cuda.synchronize()
    o = out.cpu().numpy().reshape(4, n)
    return o[0], o[1], o[2], o[3]


def _bits(x):
    return x.view(np.int64)


def test_division_sequence_is_the_ieee_quotient(gpu_ok):
    rng = np.random.default_rng(5)
    n = 1 << 20
    a = _doubles(rng, n, -300, 300)
    b = _doubles(rng, n, -300, 300)
    b[: n // 8] = _all_ones(rng, b[: n // 8])
    b[n // 8: n // 4] = rng.choice(SPECIALS, n // 8)
    # the solver's own magnitudes: densities / energies over densities, sound speeds
    a[n // 4: n // 2] = rng.uniform(-1.0e4, 1.0e4, n // 4)
    b[n // 4: n // 2] = 10.0 ** rng.uniform(-10.0, 5.0, n // 4)
    got, ieee, _, _ = _run(a, b)
    want = a / b
    assert np.array_equal(_bits(ieee), _bits(want)), "device IEEE division differs from the host"
    bad = np.flatnonzero(_bits(got) != _bits(want))
    assert bad.size == 0, [(a[i], b[i], got[i], want[i]) for i in bad[:5]]


def test_division_edges_of_the_documented_range(gpu_ok):
    """|b| at 2^-1000 and 2^1000, quotients down to 2^-1000 and up to 2^700,
    and zero numerators (a zero may come back with either sign)."""
    rng = np.random.default_rng(6)
    n = 1 << 16
    b = np.concatenate([_doubles(rng, n, -1000, -990), _doubles(rng, n, 990, 999)])
    q = np.concatenate([_doubles(rng, n, 600, 699), _doubles(rng, n, -990, -900)])
    a = q * b
    a[(np.abs(a) < 2.0 ** -960) | ~np.isfinite(a)] = 1.0
    a[::97] = 0.0
    keep = (np.abs(a / b) >= 2.0 ** -1000) & (np.abs(a / b) <= 2.0 ** 700) | (a == 0.0)
    a, b = a[keep], b[keep]
    got, _, _, _ = _run(a, b)
    want = a / b
    nz = a != 0.0
    assert np.array_equal(_bits(got[nz]), _bits(want[nz]))
    assert np.all(got[~nz] == 0.0)


def test_square_root_sequence_is_the_ieee_root(gpu_ok):
    rng = np.random.default_rng(7)
    n = 1 << 20
    x = np.abs(_doubles(rng, n, -760, 1020))
    x[: n // 4] = 10.0 ** rng.uniform(-20.0, 12.0, n // 4)  # c^2 clamped to >= 1e-20
    x[n // 4: n // 2] = _all_ones(rng, x[n // 4: n // 2])
    x[:16] = [1.0e-20, 2.0 ** -767, 1.0, 2.0, 4.0, 0.25, 1.0e12, 2.0 ** 1000,
              np.nextafter(1.0, 2.0), np.nextafter(1.0, 0.0), 3.0, 5.0, 1.0e-10, 6000.0,
              np.finfo(np.float64).max, 2.0 ** -766]
    _, _, got, ieee = _run(np.ones_like(x), x)
    want = np.sqrt(x)
    assert np.array_equal(_bits(ieee), _bits(want)), "device IEEE sqrt differs from the host"
    bad = np.flatnonzero(_bits(got) != _bits(want))
    assert bad.size == 0, [(x[i], got[i], want[i]) for i in bad[:5]]


def test_selftest_rejects_bad_arguments(gpu_ok):
    from wamp import _lib

    lib = _lib.load()
    assert lib.wamp_selftest_arith(None, None, None, 0, None) == 0
    assert lib.wamp_selftest_arith(None, None, None, 4, None) != 0
    assert lib.wamp_selftest_arith(None, None, None, -1, None) != 0


def test_elementary_functions_are_the_host_build_bit_for_bit(gpu_ok):
    """pow and tanh of the FP64 kernels (wamp_math.h) against the same source
    compiled by gcc into the oracle build liborc_wm.so: identical bits over
    the step's argument ranges and beyond (tests/test_math.py checks that
    source against 60-digit arithmetic)."""
    import torch

    from oracle import oracle as O
    from wamp import _lib

    lib = _lib.load()
    rng = np.random.default_rng(9)
    n = 200_000
    x = np.concatenate([10.0 ** rng.uniform(-20, 6, n // 2), rng.uniform(-30.0, 30.0, n // 4),
                        rng.uniform(-0.9, 0.9, n // 4)])
    y = np.concatenate([rng.uniform(1.0, 7.0, n // 2), rng.uniform(-3.0, 3.0, n // 2)])
    x[:8] = [0.0, -0.0, 1e-300, 5e-324, 1.0, 0.9, -0.9, 1e-20]
    out = torch.zeros(2 * n, dtype=torch.float64, device="cuda")
    dx, dy = torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda()
    _lib.check(lib.wamp_selftest_elementary(dx.data_ptr(), dy.data_ptr(), out.data_ptr(), n,
                                            torch.cuda.current_stream().cuda_stream),
               "wamp_selftest_elementary")
    got = out.cpu().numpy()
    W = O.lib(shared_math=True)
    px = np.abs(x)   # pow's domain in the step: positive bases
    want_pow = np.array([W.orc_pow(float(a), float(b)) for a, b in zip(px, y)])
    dpx = torch.from_numpy(px).cuda()   # held until the kernel has run
    _lib.check(lib.wamp_selftest_elementary(dpx.data_ptr(), dy.data_ptr(), out.data_ptr(), n,
                                            torch.cuda.current_stream().cuda_stream),
               "wamp_selftest_elementary")
    got_pow = out[:n].cpu().numpy()
    want_tanh = np.array([W.orc_tanh(float(a)) for a in x])
    assert np.array_equal(got_pow.view(np.uint64), want_pow.view(np.uint64))
    assert np.array_equal(got[n:].view(np.uint64), want_tanh.view(np.uint64))
