/* Body of the CPU oracle for the kinematic sum, instantiated twice by
 * kin_sum.c: once with REAL = double (the reference's own precision) and once
 * with REAL = long double (x87 80-bit, 64-bit significand) as a
 * higher-precision truth.
 *
 * Semantics follow compute_scattering_cffi, Beam.py:663-816, term by term:
 *   per pixel p:  R0 = |T_p|, Q_cut = 0.5 |(Q_x, Q_y)| from the right/up
 *                 neighbours (inner neighbour at the last column/row)
 *                 (Beam.py:718-746);
 *   per pair:     d = T_p - r_j, r = |d|, dotv = d_x / r,
 *                 Q = k sqrt(max(0, 2 (1 - dotv))),
 *                 f0 = c + sum_i a_i exp(-b_i s^2), s = Q * 0.25e-10 / pi
 *                 (Beam.py:669-679),
 *                 S = f0 + f' + i f''  (minus f0(0) + f' + i f'' when
 *                 remove_forward and Q < Q_cut[p]),
 *                 t = amp_j * S,
 *                 phase = fmod(k (x_j + r), 2 pi),
 *                 val = t * exp(i phase) * r_e * [R0 / r] * [sqrt(clip(P))],
 *                 P = 0.5 (1 - p)(1 + dotv^2) + p (1 - (e . d / r)^2)
 *                 (Beam.py:748-807).
 * Pairs with r == 0 are skipped (Beam.py:760).
 *
 * Differences from the reference, on purpose:
 *   - the loop is pixel-outer and parallel over pixels (OpenMP), each pixel
 *     summing every atom in index order, so the result does not depend on the
 *     thread count;
 *   - outputs are REAL accumulators added into out_r/out_i as double,
 *     instead of one float32 cast per call.
 */
#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(name, SUFFIX)

static REAL FN(f0_value)(REAL Q, const float* prm)
{
    const REAL s = ((REAL)0.25e-10 / PI_R) * Q;
    const REAL ss = s * s;
    REAL f = (REAL)prm[5];
    for (int i = 0; i < 5; ++i)
        f += (REAL)prm[i] * EXP_R(-(REAL)prm[6 + i] * ss);
    return f;
}

int FN(oracle_kin_sum)(
    int atom_count, const double* positions, const float* f0_params,
    const float* f0_zero, int remove_forward, const float* s_anom_real,
    const float* s_anom_imag, const float* amp_r, const float* amp_i,
    int Ny, int Nz, const double* cx, const double* cy, const double* cz,
    double k_val_d, int apply_pol, double pol_deg, double pol_ex,
    double pol_ey, double pol_ez, int apply_decay, double* out_r,
    double* out_i, int n_threads)
{
    const REAL TWO_PI = (REAL)2 * PI_R;
    const REAL rE = (REAL)2.81794092e-15;
    const REAL k_val = (REAL)k_val_d;
    const int npix = Ny * Nz;
    if (atom_count < 0 || Ny <= 0 || Nz <= 0) return 2;
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#pragma omp parallel for schedule(dynamic, 16)
#endif
    for (int p = 0; p < npix; ++p) {
        const int ix = p % Ny, iy = p / Ny;
        const REAL tx = cx[p], ty = cy[p], tz = cz[p];
        const REAL R0 = SQRT_R(tx * tx + ty * ty + tz * tz);
        REAL ux = 0, uy = 0, uz = 0;
        if (R0 > 0) { ux = tx / R0; uy = ty / R0; uz = tz / R0; }
        REAL qcut;
        {
            const int nr = (ix + 1 < Ny) ? (p + 1) : ((ix > 0) ? (p - 1) : p);
            const int nu = (iy + 1 < Nz) ? (p + Ny) : ((iy > 0) ? (p - Ny) : p);
            REAL q2[2];
            const int nb[2] = {nr, nu};
            for (int t = 0; t < 2; ++t) {
                const REAL rx = cx[nb[t]], ry = cy[nb[t]], rz = cz[nb[t]];
                const REAL Rr = SQRT_R(rx * rx + ry * ry + rz * rz);
                REAL cd = (Rr > 0) ? (ux * rx + uy * ry + uz * rz) / Rr : (REAL)1;
                if (cd > 1) cd = 1;
                if (cd < -1) cd = -1;
                REAL a = 2 * (1 - cd);
                if (a < 0) a = 0;
                q2[t] = k_val * SQRT_R(a);
            }
            qcut = (REAL)0.5 * SQRT_R(q2[0] * q2[0] + q2[1] * q2[1]);
        }
        REAL acc_r = 0, acc_i = 0;
        for (int a = 0; a < atom_count; ++a) {
            const REAL ax = positions[3 * a + 0];
            const REAL ay = positions[3 * a + 1];
            const REAL az = positions[3 * a + 2];
            const REAL dx = tx - ax, dy = ty - ay, dz = tz - az;
            const REAL r = SQRT_R(dx * dx + dy * dy + dz * dz);
            if (r == 0) continue;
            const REAL dotv = dx / r;
            REAL arg = 2 * (1 - dotv);
            if (arg < 0) arg = 0;
            const REAL Q = k_val * SQRT_R(arg);
            const REAL f0 = FN(f0_value)(Q, &f0_params[11 * a]);
            REAL s_re = f0 + (REAL)s_anom_real[a];
            REAL s_im = (REAL)s_anom_imag[a];
            if (remove_forward && Q < qcut) {
                s_re -= ((REAL)f0_zero[a] + (REAL)s_anom_real[a]);
                s_im -= (REAL)s_anom_imag[a];
            }
            const REAL ar = amp_r[a], ai = amp_i[a];
            const REAL t_re = ar * s_re - ai * s_im;
            const REAL t_im = ar * s_im + ai * s_re;
            const REAL ph = FMOD_R(k_val * (ax + r), TWO_PI);
            const REAL c = COS_R(ph), s = SIN_R(ph);
            const REAL v_re = t_re * c - t_im * s;
            const REAL v_im = t_re * s + t_im * c;
            REAL scale = rE;
            if (apply_decay && R0 > 0) scale *= R0 / r;
            if (apply_pol) {
                const REAL er = ((REAL)pol_ex * dx + (REAL)pol_ey * dy + (REAL)pol_ez * dz) / r;
                REAL P = (REAL)0.5 * (1 - (REAL)pol_deg) * (1 + dotv * dotv)
                       + (REAL)pol_deg * (1 - er * er);
                if (P < 0) P = 0;
                if (P > 1) P = 1;
                scale *= SQRT_R(P);
            }
            acc_r += v_re * scale;
            acc_i += v_im * scale;
        }
        out_r[p] += (double)acc_r;
        out_i[p] += (double)acc_i;
    }
    return 0;
}

#undef FN
#undef CAT
#undef CAT2
