"""GPU: the allocation-free expecPauliString against the allocating one.

fsa_densitymatrix_expecPauliString_noalloc evaluates Tr(H rho) with its term
list and partial sums in a static device array instead of memory the library
allocates.  It runs the same reduction kernels with the same grid, so its
result must equal the allocating call's bit for bit.  Two calls in a row, the
second on a smaller input (fewer terms, fewer qubits), show that what the first
call left in the shared array does not leak into the second.
"""
import ctypes as C

import numpy as np
import pytest

import __graft_entry__ as ge

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fsa():
    return ge.load_package()


def noalloc_expec(fsa, rho, coeffs, paulis):
    f = fsa.lib().fsa_densitymatrix_expecPauliString_noalloc
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_uint, C.c_void_p, C.c_void_p]
    f.restype = C.c_int
    c = np.ascontiguousarray(coeffs, dtype=np.float64)
    p = np.ascontiguousarray(paulis, dtype=np.uint32)
    out = np.zeros(2)
    code = f(rho.handle, c.ctypes.data_as(C.c_void_p), len(c), p.ctypes.data_as(C.c_void_p),
             out.ctypes.data_as(C.c_void_p))
    if code:
        raise fsa.FsaError(code, fsa.lib().fsa_last_error().decode())
    return complex(out[0], out[1])


def case(fsa, N, T, seed, precision):
    rng = np.random.default_rng(seed)
    rho = fsa.DensityMatrix(N, precision=precision)
    v = rng.normal(size=1 << 2 * N) + 1j * rng.normal(size=1 << 2 * N)
    rho.upload(v / np.linalg.norm(v))
    rho.sync()
    coeffs = rng.normal(size=T)
    paulis = rng.integers(0, 4, T * N)
    return rho, coeffs, paulis


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_noalloc_expec_matches_allocating_call(fsa, precision):
    for N, T, seed in ((12, 64, 1), (7, 5, 2)):        # the second call is smaller
        rho, coeffs, paulis = case(fsa, N, T, seed, precision)
        got = noalloc_expec(fsa, rho, coeffs, paulis)
        want = fsa.distributed_densitymatrix_expecPauliString(rho, coeffs, paulis)
        print(f"{precision} N={N} T={T}: noalloc {got} allocating {want}")
        assert got == want, (N, T, got, want)
        rho.close()


def test_noalloc_expec_rejects_what_it_cannot_do(fsa):
    rho, coeffs, paulis = case(fsa, 4, 3, 3, "fp64")
    # queued operations would need a flush, which may allocate
    fsa.distributed_densitymatrix_damping(rho, 1, 0.2)
    with pytest.raises(fsa.FsaError) as e:
        noalloc_expec(fsa, rho, coeffs, paulis)
    assert e.value.code == 1
    rho.sync()
    assert noalloc_expec(fsa, rho, coeffs, paulis) == fsa.distributed_densitymatrix_expecPauliString(rho, coeffs, paulis)
    # more terms than the scratch array holds: a clear size error
    big_c = np.ones(120000)
    big_p = np.zeros(120000 * 4, dtype=np.uint32)
    with pytest.raises(fsa.FsaError) as e:
        noalloc_expec(fsa, rho, big_c, big_p)
    assert e.value.code == 4
    sv = fsa.StateVector(3)
    with pytest.raises(fsa.FsaError) as e:
        noalloc_expec(fsa, sv, [1.0], [3, 0, 0])
    assert e.value.code == 7
