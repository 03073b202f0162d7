// types.hpp -- drop-in for the reference's src/types.hpp (same names and
// conventions): Pauli codes, precision typedefs, complex/array/matrix types
// and the small host-side dense-matrix helpers gate construction uses.
#ifndef TYPES_HPP
#define TYPES_HPP

#include <assert.h>
#include <complex>
#include <vector>

enum PauliOperator { I, X, Y, Z };               // types.hpp:16-18

typedef double Real;                             // types.hpp:26-28
typedef unsigned int Nat;
typedef long long unsigned int Index;

typedef std::complex<Real> Amp;                  // types.hpp:37-42
typedef std::vector<Amp> AmpArray;
typedef std::vector<AmpArray> AmpMatrix;
typedef std::vector<Nat> NatArray;
typedef std::vector<AmpMatrix> MatrixArray;
typedef std::vector<Real> RealArray;

static inline AmpMatrix getZeroMatrix(Index dim) {
    assert(dim > 0);
    return AmpMatrix(dim, AmpArray(dim, 0));
}

static inline AmpMatrix getIdentityMatrix(Index dim) {
    AmpMatrix m = getZeroMatrix(dim);
    for (Index r = 0; r < dim; r++) m[r][r] = 1;
    return m;
}

static inline AmpMatrix getConjugateMatrix(AmpMatrix m) {
    for (auto& row : m)
        for (auto& v : row) v = std::conj(v);
    return m;
}

static inline AmpMatrix getDaggerMatrix(const AmpMatrix& m) {
    AmpMatrix d = getZeroMatrix(m.size());
    for (Index r = 0; r < m.size(); r++)
        for (Index c = 0; c < m.size(); c++) d[r][c] = std::conj(m[c][r]);
    return d;
}

static inline AmpMatrix operator*(const Amp& f, const AmpMatrix& m) {
    AmpMatrix out = m;
    for (auto& row : out)
        for (auto& v : row) v *= f;
    return out;
}

static inline AmpMatrix operator*(const AmpMatrix& a, const AmpMatrix& b) {
    assert(a.size() == b.size());
    AmpMatrix p = getZeroMatrix(a.size());
    for (Index r = 0; r < a.size(); r++)
        for (Index k = 0; k < a.size(); k++)
            for (Index c = 0; c < a.size(); c++) p[r][c] += a[r][k] * b[k][c];
    return p;
}

static inline AmpMatrix operator+(const AmpMatrix& a, const AmpMatrix& b) {
    assert(a.size() == b.size());
    AmpMatrix s = a;
    for (Index r = 0; r < a.size(); r++)
        for (Index c = 0; c < a.size(); c++) s[r][c] += b[r][c];
    return s;
}

// Kronecker product with b on the low (fast) index, as types.hpp:128-138
static inline AmpMatrix operator%(const AmpMatrix& a, const AmpMatrix& b) {
    const Index da = a.size(), db = b.size();
    AmpMatrix p = getZeroMatrix(da * db);
    for (Index i = 0; i < da; i++)
        for (Index j = 0; j < da; j++)
            for (Index r = 0; r < db; r++)
                for (Index c = 0; c < db; c++) p[r + db * i][c + db * j] = a[i][j] * b[r][c];
    return p;
}

static inline AmpArray operator*(const AmpMatrix& m, const AmpArray& v) {
    assert(m.size() == v.size());
    AmpArray p(v.size(), 0);
    for (Index r = 0; r < v.size(); r++)
        for (Index c = 0; c < v.size(); c++) p[r] += m[r][c] * v[c];
    return p;
}

#endif  // TYPES_HPP
