"""MRG32k3a pins: oracle, product host code and an exact-integer Python
generator against the reference's own rngstream.cpp / ParallelRng.cpp
(oracle/_ref), and against what the algebra of the generator fixes."""
import ctypes as C

import numpy as np
import pytest

from casim_amd import abi, rng

SEEDS = [rng.DEFAULT_SEED, (1, 2, 3, 4, 5, 6), (4294967086, 1, 0, 4294944442, 0, 7), (0, 0, 1, 0, 0, 1)]


def ref_stream(ref_lib, seed, k, n):
    out = np.empty(n)
    st = (C.c_ulong * 6)()
    assert ref_lib.ref_stream_draws((C.c_ulong * 6)(*seed), k, n, out.ctypes.data_as(C.POINTER(C.c_double)), st) == 0
    return out, np.array(list(st), dtype=np.uint64)


@pytest.mark.parametrize("seed", SEEDS)
@pytest.mark.parametrize("k", [0, 1, 2, 7, 31])
def test_streams_match_reference(ref_lib, oracle_mod, seed, k):
    want, want_state = ref_stream(ref_lib, seed, k, 300)
    start = rng.stream_states(seed, 1, first=k)[0]
    assert np.array_equal(oracle_mod.stream_states(np.array(seed, dtype=np.uint64), 1, first=k)[0], start)
    got, st = oracle_mod.rng_draws(start, 300)
    assert np.array_equal(got, want)
    assert np.array_equal(st, want_state)
    py = rng.Stream(start)
    assert np.array_equal(np.array([py.rand_u01() for _ in range(300)]), want)
    lib = abi.load()
    s = start.copy()
    out = np.empty(300)
    abi.check(lib.casim_rng_draws(abi.uptr(s), 300, abi.dptr(out)))
    assert np.array_equal(out, want) and np.array_equal(s, want_state)


def test_product_stream_states_and_advance(oracle_mod):
    lib = abi.load()
    for seed in SEEDS:
        a = abi.rng_stream_states(seed, 40)
        assert np.array_equal(a, rng.stream_states(seed, 40))
        s = a[3].copy()
        abi.check(lib.casim_rng_advance(abi.uptr(s), 1234))
        _, st = oracle_mod.rng_draws(a[3], 1234)
        assert np.array_equal(s, st)


def test_parallel_rng_uses_stream_zero(ref_lib, oracle_mod):
    """A JSON "Seed" goes to ParallelRng(seed): the master thread draws from
    stream 0, i.e. the seed itself (ParallelRng.cpp:16-21,33-35)."""
    seed = (11, 22, 33, 44, 55, 66)
    out = np.empty(50)
    ref_lib.ref_parallel_rng_draws((C.c_ulong * 6)(*seed), 50, out.ctypes.data_as(C.POINTER(C.c_double)))
    got, _ = oracle_mod.rng_draws(np.array(seed, dtype=np.uint64), 50)
    assert np.array_equal(got, out)


def test_substream(ref_lib, oracle_mod):
    for k in (0, 3):
        st = (C.c_ulong * 6)()
        ref_lib.ref_stream_substream((C.c_ulong * 6)(*rng.DEFAULT_SEED), k, st)
        start = rng.stream_states(rng.DEFAULT_SEED, 1, first=k)[0]
        assert np.array_equal(rng.next_substream(start), np.array(list(st), dtype=np.uint64))
        s2 = start.copy()
        oracle_mod.load().oracle_rng_substream(oracle_mod.u(s2))
        assert np.array_equal(s2, np.array(list(st), dtype=np.uint64))


def test_jump_matrices_are_powers(oracle_mod):
    """The oracle's A^(2^e) (double arithmetic) equals exact integer powers."""
    for which, (A, m) in ((1, (rng.A1, rng.M1)), (2, (rng.A2, rng.M2))):
        for e in (1, 5, 76, 127):
            out = np.empty(9)
            oracle_mod.load().oracle_rng_jump_matrix(which, e, oracle_mod.d(out))
            want = np.array(rng.matpow(A, 1 << e, m), dtype=np.float64).reshape(-1)
            assert np.array_equal(out, want), (which, e)


def test_reference_advance_state(ref_lib):
    """AdvanceState(127, 0) from stream k's start lands on stream k+1's start."""
    st = (C.c_ulong * 6)()
    ref_lib.ref_stream_advance((C.c_ulong * 6)(*rng.DEFAULT_SEED), 127, 0, st)
    assert np.array_equal(np.array(list(st), dtype=np.uint64), rng.stream_states(rng.DEFAULT_SEED, 2)[1])


def test_draw_properties():
    """U01d values lie in [0, 1) and are 53-bit: u = U01 + U01 2^-24 (mod 1)."""
    s = rng.Stream(rng.DEFAULT_SEED)
    u = np.array([s.rand_u01() for _ in range(20000)])
    assert u.min() >= 0.0 and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 0.01
    assert abs(u.var() - 1 / 12) < 0.003


def test_seed_check():
    assert not rng.check_seed((0, 0, 0, 1, 1, 1))
    assert not rng.check_seed((4294967087, 1, 1, 1, 1, 1))
    assert not rng.check_seed((1, 1, 1, 4294944443, 1, 1))
    lib = abi.load()
    bad = np.array([0, 0, 0, 1, 2, 3], dtype=np.uint64)
    out = np.empty(2, dtype=np.uint64)
    assert lib.casim_rng_stream_states(abi.uptr(bad), 0, 1, abi.uptr(np.zeros(6, dtype=np.uint64))) == abi_invalid()


def abi_invalid():
    return -1
