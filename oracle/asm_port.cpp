// Oracle: the reference's per-tet Nedelec-2 assembly kernel restated in C++
// with OpenMP over tets.
//
// TEST INFRASTRUCTURE ONLY (see oracle/__init__.py): the checker and the
// timed CPU baseline, never the product path.
//
// Restates ASM/curlcurl.py (ASM/ = src/emerge/_emerge/physics/microwave/
// assembly/): matrix_builder (:419-466, numba prange over tets writing the
// COO block of tet t at 400 t: rows[p + 20 i + j] = dof[i], cols = dof[j]),
// ned2_tet_stiff_mass (:180-412: edge-edge :213-264, edge-face :266-317,
// mirrored blocks :321-328, face-face :330-407), tet_coefficients_bcd
// (:99-138), the volume-coefficient table (:64-92), local_mapping
// (E/mth/optimized.py:293-327), compute_distances (:330-338) and matinv
// (:342-370).  The operation structure follows oracle/curlcurl.py, which
// tests/test_oracle_element.py pins against brute-force quadrature;
// tests/test_oracle_port.py checks this port against it entry by entry.
#include <cmath>
#include <complex>
#include <cstdint>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

using cd = std::complex<double>;

double g_vc[5][5][5][5];  // volume_coeff(a, b, c, d), letters 1..4, 0 = no factor

double volume_coeff(int a, int b, int c, int d) {
    static const double fac[10] = {1, 1, 2, 6, 24, 120, 720, 5040, 40320, 362880};
    int k[7] = {0, 0, 0, 0, 0, 0, 0};
    k[a]++;
    k[b]++;
    k[c]++;
    k[d]++;
    double num = 1;
    int tot = 0;
    for (int i = 1; i < 7; ++i) {
        num *= fac[k[i]];
        tot += k[i];
    }
    return num / fac[tot + 3];
}

struct Init {
    Init() {
        for (int a = 0; a < 5; ++a)
            for (int b = 0; b < 5; ++b)
                for (int c = 0; c < 5; ++c)
                    for (int d = 0; d < 5; ++d) g_vc[a][b][c][d] = volume_coeff(a, b, c, d);
    }
} g_init;

inline void mv(const cd M[3][3], const cd v[3], cd out[3]) {
    for (int i = 0; i < 3; ++i) out[i] = M[i][0] * v[0] + M[i][1] * v[1] + M[i][2] * v[2];
}
inline cd dot(const cd a[3], const cd b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline void cross(const cd a[3], const cd b[3], cd out[3]) {
    out[0] = a[1] * b[2] - a[2] * b[1];
    out[1] = a[2] * b[0] - a[0] * b[2];
    out[2] = a[0] * b[1] - a[1] * b[0];
}

void matinv(const cd M[3][3], cd out[3][3]) {
    const bool diag = M[0][1] == 0.0 && M[0][2] == 0.0 && M[1][0] == 0.0 && M[1][2] == 0.0 && M[2][0] == 0.0 &&
                      M[2][1] == 0.0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) out[i][j] = 0.0;
    if (diag) {
        for (int i = 0; i < 3; ++i) out[i][i] = 1.0 / M[i][i];
        return;
    }
    const cd det = M[0][0] * M[1][1] * M[2][2] - M[0][0] * M[1][2] * M[2][1] - M[0][1] * M[1][0] * M[2][2] +
                   M[0][1] * M[1][2] * M[2][0] + M[0][2] * M[1][0] * M[2][1] - M[0][2] * M[1][1] * M[2][0];
    out[0][0] = (M[1][1] * M[2][2] - M[1][2] * M[2][1]) / det;
    out[0][1] = (-M[0][1] * M[2][2] + M[0][2] * M[2][1]) / det;
    out[0][2] = (M[0][1] * M[1][2] - M[0][2] * M[1][1]) / det;
    out[1][0] = (-M[1][0] * M[2][2] + M[1][2] * M[2][0]) / det;
    out[1][1] = (M[0][0] * M[2][2] - M[0][2] * M[2][0]) / det;
    out[1][2] = (-M[0][0] * M[1][2] + M[0][2] * M[1][0]) / det;
    out[2][0] = (M[1][0] * M[2][1] - M[1][1] * M[2][0]) / det;
    out[2][1] = (-M[0][0] * M[2][1] + M[0][1] * M[2][0]) / det;
    out[2][2] = (M[0][0] * M[1][1] - M[0][1] * M[1][0]) / det;
}

// cofactors b, c, d and V = |det| / 6 (ASM/curlcurl.py:99-138)
void bcd(const double x[4], const double y[4], const double z[4], double b[4], double c[4], double d[4], double& V) {
    const double x1 = x[0], x2 = x[1], x3 = x[2], x4 = x[3];
    const double y1 = y[0], y2 = y[1], y3 = y[2], y4 = y[3];
    const double z1 = z[0], z2 = z[1], z3 = z[2], z4 = z[3];
    V = std::fabs(-x1 * y2 * z3 / 6 + x1 * y2 * z4 / 6 + x1 * y3 * z2 / 6 - x1 * y3 * z4 / 6 - x1 * y4 * z2 / 6 +
                  x1 * y4 * z3 / 6 + x2 * y1 * z3 / 6 - x2 * y1 * z4 / 6 - x2 * y3 * z1 / 6 + x2 * y3 * z4 / 6 +
                  x2 * y4 * z1 / 6 - x2 * y4 * z3 / 6 - x3 * y1 * z2 / 6 + x3 * y1 * z4 / 6 + x3 * y2 * z1 / 6 -
                  x3 * y2 * z4 / 6 - x3 * y4 * z1 / 6 + x3 * y4 * z2 / 6 + x4 * y1 * z2 / 6 - x4 * y1 * z3 / 6 -
                  x4 * y2 * z1 / 6 + x4 * y2 * z3 / 6 + x4 * y3 * z1 / 6 - x4 * y3 * z2 / 6);
    b[0] = -y2 * z3 + y2 * z4 + y3 * z2 - y3 * z4 - y4 * z2 + y4 * z3;
    b[1] = y1 * z3 - y1 * z4 - y3 * z1 + y3 * z4 + y4 * z1 - y4 * z3;
    b[2] = -y1 * z2 + y1 * z4 + y2 * z1 - y2 * z4 - y4 * z1 + y4 * z2;
    b[3] = y1 * z2 - y1 * z3 - y2 * z1 + y2 * z3 + y3 * z1 - y3 * z2;
    c[0] = x2 * z3 - x2 * z4 - x3 * z2 + x3 * z4 + x4 * z2 - x4 * z3;
    c[1] = -x1 * z3 + x1 * z4 + x3 * z1 - x3 * z4 - x4 * z1 + x4 * z3;
    c[2] = x1 * z2 - x1 * z4 - x2 * z1 + x2 * z4 + x4 * z1 - x4 * z2;
    c[3] = -x1 * z2 + x1 * z3 + x2 * z1 - x2 * z3 - x3 * z1 + x3 * z2;
    d[0] = -x2 * y3 + x2 * y4 + x3 * y2 - x3 * y4 - x4 * y2 + x4 * y3;
    d[1] = x1 * y3 - x1 * y4 - x3 * y1 + x3 * y4 + x4 * y1 - x4 * y3;
    d[2] = -x1 * y2 + x1 * y4 + x2 * y1 - x2 * y4 - x4 * y1 + x4 * y2;
    d[3] = x1 * y2 - x1 * y3 - x2 * y1 + x2 * y3 + x3 * y1 - x3 * y2;
}

// tet-local index of a global vertex id: first match (E/mth/optimized.py:293-327)
inline int local_index(const int64_t vid[4], int64_t g) {
    for (int k = 0; k < 4; ++k)
        if (vid[k] == g) return k;
    return 0;
}

void tet_matrices(const double xs[4], const double ys[4], const double zs[4], const double L[6], const int lem[2][6],
                  const int ltm[3][4], const cd Ms[3][3], const cd Mm[3][3], cd Dm[20][20], cd Fm[20][20]) {
    double b[4], c[4], d[4], V;
    bcd(xs, ys, zs, b, c, d, V);
    double Ds[4][4];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            Ds[i][j] = std::sqrt((xs[i] - xs[j]) * (xs[i] - xs[j]) + (ys[i] - ys[j]) * (ys[i] - ys[j]) +
                                 (zs[i] - zs[j]) * (zs[i] - zs[j]));
    cd GL[4][3];
    for (int i = 0; i < 4; ++i) {
        GL[i][0] = b[i];
        GL[i][1] = c[i];
        GL[i][2] = d[i];
    }
    const double V6 = 6 * V;
    const double KA = 1.0 / std::pow(V6, 4), KB = 1.0 / (V6 * V6);
    auto VC = [&](int a, int b_, int c_, int d_) { return g_vc[a][b_][c_][d_] * V6; };
    for (int i = 0; i < 20; ++i)
        for (int j = 0; j < 20; ++j) Dm[i][j] = Fm[i][j] = 0.0;

    for (int ei = 0; ei < 6; ++ei) {
        const int ei1 = lem[0][ei], ei2 = lem[1][ei];
        const cd* GA = GL[ei1];
        const cd* GB = GL[ei2];
        const int A = ei1 + 1, B = ei2 + 1;
        const double L1 = L[ei];
        cd cab[3];
        cross(GA, GB, cab);
        for (int ej = 0; ej < 6; ++ej) {
            const int ej1 = lem[0][ej], ej2 = lem[1][ej];
            const int C = ej1 + 1, D = ej2 + 1;
            const cd* GC = GL[ej1];
            const cd* GD = GL[ej2];
            const double L2 = L[ej];
            cd mGC[3], mGD[3], ccd[3], mcd[3];
            mv(Mm, GC, mGC);
            mv(Mm, GD, mGD);
            const cd g_ac = dot(GA, mGC), g_ad = dot(GA, mGD), g_bc = dot(GB, mGC), g_bd = dot(GB, mGD);
            const double L12 = L1 * L2;
            cross(GC, GD, ccd);
            mv(Ms, ccd, mcd);
            const cd factor = L12 * 9.0 * dot(cab, mcd);
            Dm[ei][ej] = factor * VC(A, C, 0, 0);
            Dm[ei][ej + 10] = factor * VC(A, D, 0, 0);
            Dm[ei + 10][ej] = factor * VC(B, C, 0, 0);
            Dm[ei + 10][ej + 10] = factor * VC(B, D, 0, 0);
            Fm[ei][ej] = L12 * (VC(A, B, C, D) * g_ac - VC(A, B, C, C) * g_ad - VC(A, A, C, D) * g_bc +
                                VC(A, A, C, C) * g_bd);
            Fm[ei][ej + 10] = L12 * (VC(A, B, D, D) * g_ac - VC(A, B, C, D) * g_ad - VC(A, A, D, D) * g_bc +
                                     VC(A, A, C, D) * g_bd);
            Fm[ei + 10][ej] = L12 * (VC(B, B, C, D) * g_ac - VC(B, B, C, C) * g_ad - VC(A, B, C, D) * g_bc +
                                     VC(A, B, C, C) * g_bd);
            Fm[ei + 10][ej + 10] = L12 * (VC(B, B, D, D) * g_ac - VC(B, B, C, D) * g_ad - VC(A, B, D, D) * g_bc +
                                          VC(A, B, C, D) * g_bd);
        }
        for (int ej = 0; ej < 4; ++ej) {
            const int ej1 = ltm[0][ej], ej2 = ltm[1][ej], fj = ltm[2][ej];
            const int C = ej1 + 1, D = ej2 + 1, F = fj + 1;
            const cd* GC = GL[ej1];
            const cd* GD = GL[ej2];
            const cd* GF = GL[fj];
            const double Lab2 = Ds[ej1][ej2], Lac2 = Ds[ej1][fj];
            cd t1[3], t2[3];
            cross(GC, GF, t1);
            mv(Ms, t1, t2);
            const cd h_df = dot(cab, t2);
            cross(GD, GF, t1);
            mv(Ms, t1, t2);
            const cd h_cd = dot(cab, t2);
            cross(GC, GD, t1);
            mv(Ms, t1, t2);
            const cd h_cf = dot(cab, t2);
            cd mGF[3], mGC[3], mGD[3];
            mv(Mm, GF, mGF);
            mv(Mm, GC, mGC);
            mv(Mm, GD, mGD);
            const cd g_af = dot(GA, mGF), g_ac = dot(GA, mGC), g_ad = dot(GA, mGD);
            const cd g_bf = dot(GB, mGF), g_bc = dot(GB, mGC), g_bd = dot(GB, mGD);
            Dm[ei][ej + 6] = L1 * Lac2 * (-6.0 * VC(A, D, 0, 0) * h_df - 3.0 * VC(A, C, 0, 0) * h_cd -
                                          3.0 * VC(A, F, 0, 0) * h_cf);
            Dm[ei][ej + 16] = L1 * Lab2 * (6.0 * VC(A, F, 0, 0) * h_cf + 3.0 * VC(A, D, 0, 0) * h_df -
                                           3.0 * VC(A, C, 0, 0) * h_cd);
            Dm[ei + 10][ej + 6] = L1 * Lac2 * (-6.0 * VC(B, D, 0, 0) * h_df - 3.0 * VC(B, C, 0, 0) * h_cd -
                                               3.0 * VC(B, F, 0, 0) * h_cf);
            Dm[ei + 10][ej + 16] = L1 * Lab2 * (6.0 * VC(B, F, 0, 0) * h_cf + 3.0 * VC(B, D, 0, 0) * h_df -
                                                3.0 * VC(B, C, 0, 0) * h_cd);
            Fm[ei][ej + 6] = L1 * Lac2 * (VC(A, B, C, D) * g_af - VC(A, B, D, F) * g_ac - VC(A, A, C, D) * g_bf +
                                          VC(A, A, D, F) * g_bc);
            Fm[ei][ej + 16] = L1 * Lab2 * (VC(A, B, D, F) * g_ac - VC(A, B, F, C) * g_ad - VC(A, A, D, F) * g_bc +
                                           VC(A, A, C, F) * g_bd);
            Fm[ei + 10][ej + 6] = L1 * Lac2 * (VC(B, B, C, D) * g_af - VC(B, B, D, F) * g_ac -
                                               VC(A, B, C, D) * g_bf + VC(A, B, D, F) * g_bc);
            Fm[ei + 10][ej + 16] = L1 * Lab2 * (VC(B, B, D, F) * g_ac - VC(B, B, F, C) * g_ad -
                                                VC(A, B, D, F) * g_bc + VC(A, B, F, C) * g_bd);
        }
    }
    // mirrored blocks (ASM/curlcurl.py:321-328)
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 6; ++j) {
            Dm[6 + i][j] = Dm[j][6 + i];
            Dm[16 + i][j] = Dm[j][16 + i];
            Dm[6 + i][10 + j] = Dm[10 + j][6 + i];
            Dm[16 + i][10 + j] = Dm[10 + j][16 + i];
            Fm[6 + i][j] = Fm[j][6 + i];
            Fm[16 + i][j] = Fm[j][16 + i];
            Fm[6 + i][10 + j] = Fm[10 + j][6 + i];
            Fm[16 + i][10 + j] = Fm[10 + j][16 + i];
        }
    for (int ei = 0; ei < 4; ++ei) {
        const int ei1 = ltm[0][ei], ei2 = ltm[1][ei], fi = ltm[2][ei];
        const int A = ei1 + 1, B = ei2 + 1, E = fi + 1;
        const cd* GA = GL[ei1];
        const cd* GB = GL[ei2];
        const cd* GE = GL[fi];
        const double Lac1 = Ds[ei1][fi], Lab1 = Ds[ei1][ei2];
        cd c_ae[3], c_be[3], c_ab[3];
        cross(GA, GE, c_ae);
        cross(GB, GE, c_be);
        cross(GA, GB, c_ab);
        for (int ej = 0; ej < 4; ++ej) {
            const int ej1 = ltm[0][ej], ej2 = ltm[1][ej], fj = ltm[2][ej];
            const int C = ej1 + 1, D = ej2 + 1, F = fj + 1;
            const cd* GC = GL[ej1];
            const cd* GD = GL[ej2];
            const cd* GF = GL[fj];
            const double Lac2 = Ds[ej1][fj], Lab2 = Ds[ej1][ej2];
            cd t[3], m_cf[3], m_df[3], m_cd[3];
            cross(GC, GF, t);
            mv(Ms, t, m_cf);
            cross(GD, GF, t);
            mv(Ms, t, m_df);
            cross(GC, GD, t);
            mv(Ms, t, m_cd);
            const cd ae_cf = dot(c_ae, m_cf), ae_df = dot(c_ae, m_df), ae_cd = dot(c_ae, m_cd);
            const cd be_cf = dot(c_be, m_cf), be_df = dot(c_be, m_df), be_cd = dot(c_be, m_cd);
            const cd ab_cf = dot(c_ab, m_cf), ab_df = dot(c_ab, m_df), ab_cd = dot(c_ab, m_cd);
            cd mGF[3], mGC[3], mGD[3];
            mv(Mm, GF, mGF);
            mv(Mm, GC, mGC);
            mv(Mm, GD, mGD);
            const cd g_ef = dot(GE, mGF), g_ec = dot(GE, mGC), g_ed = dot(GE, mGD);
            const cd g_af = dot(GA, mGF), g_ac = dot(GA, mGC), g_ad = dot(GA, mGD);
            const cd g_bf = dot(GB, mGF), g_bc = dot(GB, mGC), g_bd = dot(GB, mGD);
            const double vad = VC(A, D, 0, 0), vac = VC(A, C, 0, 0), vaf = VC(A, F, 0, 0);
            const double vbf = VC(B, F, 0, 0), vbc = VC(B, C, 0, 0), vbd = VC(B, D, 0, 0);
            const double vde = VC(E, D, 0, 0), vef = VC(E, F, 0, 0), vce = VC(E, C, 0, 0);
            const cd q1 = 2.0 * vad * be_cf + vac * be_df + vaf * be_cd;
            const cd l12 = -2.0 * vaf * be_cd - vad * be_cf + vac * be_df;
            Dm[ei + 6][ej + 6] = Lac1 * Lac2 * (4.0 * vbd * ae_cf + 2.0 * vbc * ae_df + 2.0 * vbf * ae_cd + q1 +
                                                2.0 * vde * ab_cf + vce * ab_df + vef * ab_cd);
            Dm[ei + 6][ej + 16] = Lac1 * Lab2 * (-4.0 * vbf * ae_cd - 2.0 * vbd * ae_cf + 2.0 * vbc * ae_df + l12 -
                                                 2.0 * vef * ab_cd - vde * ab_cf + vce * ab_df);
            Dm[ei + 16][ej + 6] = Lab1 * Lac2 * (-4.0 * vde * ab_cf - 2.0 * vce * ab_df - 2.0 * vef * ab_cd -
                                                 2.0 * vbd * ae_cf - vbc * ae_df - vbf * ae_cd + q1);
            Dm[ei + 16][ej + 16] = Lab1 * Lab2 * (4.0 * vef * ab_cd + 2.0 * vde * ab_cf - 2.0 * vce * ab_df +
                                                  2.0 * vbf * ae_cd + vbd * ae_cf - vbc * ae_df + l12);
            const double v_abcd = VC(A, B, C, D), v_abdf = VC(A, B, D, F);
            const double v_bcde = VC(B, C, D, F), v_bdef = VC(B, E, D, F);
            const double v_abcf = VC(A, B, F, C), v_bcef = VC(B, E, F, C);
            const double v_acde = VC(E, A, C, D), v_adef = VC(E, A, D, F);
            const double v_acef = VC(A, C, E, F);
            Fm[ei + 6][ej + 6] = Lac1 * Lac2 * (v_abcd * g_ef - v_abdf * g_ec - v_bcde * g_af + v_bdef * g_ac);
            Fm[ei + 6][ej + 16] = Lac1 * Lab2 * (v_abdf * g_ec - v_abcf * g_ed - v_bdef * g_ac + v_bcef * g_ad);
            Fm[ei + 16][ej + 6] = Lab1 * Lac2 * (v_bcde * g_af - v_bdef * g_ac - v_acde * g_bf + v_adef * g_bc);
            Fm[ei + 16][ej + 16] = Lab1 * Lab2 * (v_bdef * g_ac - v_bcef * g_ad - v_adef * g_bc + v_acef * g_bd);
        }
    }
    for (int i = 0; i < 20; ++i)
        for (int j = 0; j < 20; ++j) {
            Dm[i][j] *= KA;
            Fm[i][j] *= KB;
        }
}

}  // namespace

extern "C" {

// Threads of the following parallel regions (the reference's own default is
// NUMBA_NUM_THREADS = 4, src/emerge/__init__.py:42-48); 0 = all host threads.
void oracle_set_threads(int n) { omp_set_num_threads(n > 0 ? n : omp_get_num_procs()); }

// COO data of E (curl-curl) and B (mass) for tets [0, T): entry (i, j) of
// tet t at 400 t + 20 i + j, row tet_to_field[i, t], column tet_to_field[j, t]
// (ASM/curlcurl.py:419-466).  Arrays are the oracle's: nodes (3, V) f64,
// tets (4, T), edges (2, nE), tris (3, nF), tet_to_field (20, T),
// tet_to_edge (6, T) int64 row-major; edge_lengths (nE,); ur, er complex
// (3, 3, T) interleaved.  Returns the number of threads used.
int oracle_matrix_builder(int64_t V, int64_t T, int64_t nE, int64_t nF, const double* nodes, const int64_t* tets,
                          const int64_t* edges, const int64_t* tris, const double* edge_lengths,
                          const int64_t* tet_to_field, const int64_t* tet_to_edge, const double* ur, const double* er,
                          double* dataE, double* dataB, int64_t* rows, int64_t* cols) {
    (void)nF;
    int nthreads = 1;
#pragma omp parallel
    {
#pragma omp single
        {
#ifdef _OPENMP
            nthreads = omp_get_num_threads();
#endif
        }
#pragma omp for schedule(static)
        for (int64_t t = 0; t < T; ++t) {
            int64_t vid[4];
            double xs[4], ys[4], zs[4];
            for (int k = 0; k < 4; ++k) {
                vid[k] = tets[k * T + t];
                xs[k] = nodes[vid[k]];
                ys[k] = nodes[V + vid[k]];
                zs[k] = nodes[2 * V + vid[k]];
            }
            int lem[2][6], ltm[3][4];
            double L[6];
            for (int e = 0; e < 6; ++e) {
                const int64_t ge = tet_to_field[e * T + t];
                lem[0][e] = local_index(vid, edges[ge]);
                lem[1][e] = local_index(vid, edges[nE + ge]);
                L[e] = edge_lengths[tet_to_edge[e * T + t]];
            }
            for (int f = 0; f < 4; ++f) {
                const int64_t gt = tet_to_field[(6 + f) * T + t] - nE;
                for (int k = 0; k < 3; ++k) ltm[k][f] = local_index(vid, tris[k * nF + gt]);
            }
            cd U[3][3], Ms[3][3], Mm[3][3];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) {
                    const int64_t o = ((i * 3 + j) * T + t) * 2;
                    U[i][j] = cd(ur[o], ur[o + 1]);
                    Mm[i][j] = cd(er[o], er[o + 1]);
                }
            matinv(U, Ms);
            cd Dm[20][20], Fm[20][20];
            tet_matrices(xs, ys, zs, L, lem, ltm, Ms, Mm, Dm, Fm);
            const int64_t p0 = t * 400;
            for (int i = 0; i < 20; ++i)
                for (int j = 0; j < 20; ++j) {
                    const int64_t p = p0 + 20 * i + j;
                    rows[p] = tet_to_field[i * T + t];
                    cols[p] = tet_to_field[j * T + t];
                    dataE[2 * p] = Dm[i][j].real();
                    dataE[2 * p + 1] = Dm[i][j].imag();
                    dataB[2 * p] = Fm[i][j].real();
                    dataB[2 * p + 1] = Fm[i][j].imag();
                }
        }
    }
    return nthreads;
}

}  // extern "C"
