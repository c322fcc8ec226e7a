/*
 * ORACLE -- test infrastructure only (see orc_common.h).
 *
 * Triton shallow-water functors and step chain, restated from
 * Triton/src/triton_appfis.hpp of the reference, instantiated for REAL = double
 * (the reference dtype) and REAL = float (the FP32 bulk form).  Included twice by
 * orc_flood.c with REAL / SFX / the math macros defined.
 *
 * Layout: every field is (nx+2)*(ny+2), x fastest, ghost 1 (Array.hpp:579-590).
 */

#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)
#define FN(name) CAT(name, SFX)

#define FL_G ((REAL)9.81)
#define FL_SQRTG ((REAL)3.132091953)
#define FL_EPS12 ((REAL)1e-12)

#define AT(a, i, j) (a)[(size_t)(i) + (size_t)W * (size_t)(j)]

/* CFL functor, triton_appfis.hpp:567-586: DBL_MAX unless h > hextra, then
 * cn * min(dx/(|qx/h| + sqrt(G h)), dx/(|qy/h| + sqrt(G h))). */
static void FN(cfl)(int W, int nx, int ny, REAL dx, REAL cn, REAL hextra, const REAL* qx, const REAL* qy,
                    const REAL* h, REAL* out)
{
#pragma omp parallel for schedule(dynamic, 1)
    for (int j = 1; j <= ny; j++)
        for (int i = 1; i <= nx; i++) {
            AT(out, i, j) = RMAX;
            REAL hij = AT(h, i, j);
            if (hij > hextra) {
                REAL dtx = (dx / (FABS(AT(qx, i, j) / hij) + SQRT(FL_G * hij)));
                REAL dty = (dx / (FABS(AT(qy, i, j) / hij) + SQRT(FL_G * hij)));
                if (dtx < dty) AT(out, i, j) = cn * dtx;
                else AT(out, i, j) = cn * dty;
            }
        }
}

/* Reset functor, triton_appfis.hpp:589-605, over the full padded box (configF). */
static void FN(reset)(int W, int H, const REAL* h, REAL* sqrth, REAL* r[6])
{
#pragma omp parallel for schedule(dynamic, 1)
    for (int j = 0; j < H; j++)
        for (int i = 0; i < W; i++) {
            REAL hij = AT(h, i, j);
            REAL s = (REAL)0.0;
            if (hij > (REAL)0.0) s = SQRT(hij);
            AT(sqrth, i, j) = s;
            for (int q = 0; q < 6; q++) AT(r[q], i, j) = (REAL)0.0;
        }
}

/* Roe flux across one edge, shared body of the X functor (triton_appfis.hpp:608-824)
 * and the Y functor (827-1043).  (1) is the first cell of the edge, (2) the second;
 * (nx_, ny_) the edge normal.  Returns 0 when both cells are dry (no write at all),
 * else fills d1[3] (contribution to cell 1) and d2[3] (to cell 2). */
static int FN(roe)(REAL h1, REAL h2, REAL qx1, REAL qy1, REAL z1, REAL qx2, REAL qy2, REAL z2, REAL sqh1,
                   REAL sqh2, REAL nx_, REAL ny_, REAL dx, REAL dt, REAL hextra, REAL d1[3], REAL d2[3])
{
    if (!(h1 > (REAL)0.0 || h2 > (REAL)0.0)) return 0;
    int i;
    REAL alpha[3], beta[3], eigen[3], eigenE[3], eigev[3][3];
    REAL u1, u2, v1, v2;
    if (h1 >= hextra) { u1 = qx1 / h1; v1 = qy1 / h1; }
    else { u1 = (REAL)0.0; v1 = (REAL)0.0; }
    if (h2 >= hextra) { u2 = qx2 / h2; v2 = qy2 / h2; }
    else { u2 = (REAL)0.0; v2 = (REAL)0.0; }

    REAL h12 = (REAL)0.5 * (h1 + h2);
    REAL dh = h2 - h1;
    REAL c12 = SQRT(FL_G * h12);
    REAL u12 = (u1 * sqh1 + u2 * sqh2) / (sqh1 + sqh2);
    REAL v12 = (v1 * sqh1 + v2 * sqh2) / (sqh1 + sqh2);
    REAL un = u12 * nx_ + v12 * ny_;

    eigen[0] = un - c12;
    eigen[1] = un;
    eigen[2] = un + c12;
    for (i = 0; i < 3; i++) eigenE[i] = (REAL)0.0;

    REAL e1 = u1 * nx_ + v1 * ny_ - FL_SQRTG * sqh1;
    REAL e2 = u2 * nx_ + v2 * ny_ - FL_SQRTG * sqh2;
    if (e1 < (REAL)0.0 && e2 > (REAL)0.0) {
        eigenE[0] = eigen[0] - e1 * (e2 - eigen[0]) / (e2 - e1);
        eigen[0] = e1 * (e2 - eigen[0]) / (e2 - e1);
    }
    e1 = u1 * nx_ + v1 * ny_ + FL_SQRTG * sqh1;
    e2 = u2 * nx_ + v2 * ny_ + FL_SQRTG * sqh2;
    if (e1 < (REAL)0.0 && e2 > (REAL)0.0) {
        eigenE[2] = eigen[2] - e2 * (eigen[2] - e1) / (e2 - e1);
        eigen[2] = e2 * (eigen[2] - e1) / (e2 - e1);
    }

    eigev[0][0] = (REAL)1.0;
    eigev[0][1] = u12 - c12 * nx_;
    eigev[0][2] = v12 - c12 * ny_;
    eigev[1][0] = (REAL)0.0;
    eigev[1][1] = -c12 * ny_;
    eigev[1][2] = c12 * nx_;
    eigev[2][0] = (REAL)1.0;
    eigev[2][1] = u12 + c12 * nx_;
    eigev[2][2] = v12 + c12 * ny_;

    alpha[0] = (REAL)0.5 * (dh - (((qx2 - qx1) * nx_ + (qy2 - qy1) * ny_) - un * dh) / c12);
    alpha[1] = (((qy2 - qy1) - v12 * dh) * nx_ - ((qx2 - qx1) - u12 * dh) * ny_) / c12;
    alpha[2] = (REAL)0.5 * (dh + (((qx2 - qx1) * nx_ + (qy2 - qy1) * ny_) - un * dh) / c12);

    REAL dz = z2 - z1;
    REAL l1 = z1 + h1;
    REAL l2 = z2 + h2;
    REAL dzp = dz;
    REAL hp;
    if (dz >= (REAL)0.0) {
        hp = h1;
        if (l1 < z2) dzp = h1;
    } else {
        hp = h2;
        if (l2 < z1) dzp = -h2;
    }
    beta[0] = (REAL)0.5 * FL_G * (hp - (REAL)0.5 * FABS(dzp)) * dzp / c12;

    hp = h1 + alpha[0];
    if (eigen[0] * eigen[2] < (REAL)0.0 && hp > (REAL)0.0 && h1 > (REAL)0.0 && h2 > (REAL)0.0) {
        beta[0] = FMAX(beta[0], alpha[0] * eigen[0] - h1 * dx * (REAL)0.5 / dt);
        beta[0] = FMIN(beta[0], -alpha[2] * eigen[2] + h2 * dx * (REAL)0.5 / dt);
    }
    beta[1] = (REAL)0.0;
    beta[2] = -beta[0];

    for (i = 0; i < 3; i++) { d1[i] = (REAL)0.0; d2[i] = (REAL)0.0; }

    l1 = hp - beta[0] / eigen[0];
    l2 = hp + beta[2] / eigen[2];

    if (l2 < -FL_EPS12 && h2 < FL_EPS12) {
        d1[0] += (eigen[0] * alpha[0] - beta[0]) + (eigen[2] * alpha[2] - beta[2]);
    } else {
        if (l1 < -FL_EPS12 && h1 < FL_EPS12) {
            d2[0] += (eigen[0] * alpha[0] - beta[0]) + (eigen[2] * alpha[2] - beta[2]);
        } else {
            for (i = 0; i < 3; i++) {
                if (eigen[i] > (REAL)0.0) {
                    d2[0] += (eigen[i] * alpha[i] - beta[i]) * eigev[i][0];
                    d2[1] += (eigen[i] * alpha[i] - beta[i]) * eigev[i][1];
                    d2[2] += (eigen[i] * alpha[i] - beta[i]) * eigev[i][2];
                } else {
                    d1[0] += (eigen[i] * alpha[i] - beta[i]) * eigev[i][0];
                    d1[1] += (eigen[i] * alpha[i] - beta[i]) * eigev[i][1];
                    d1[2] += (eigen[i] * alpha[i] - beta[i]) * eigev[i][2];
                }
            }
        }
    }
    for (i = 0; i < 3; i++) {
        if (eigenE[i] > (REAL)0.0) {
            d2[0] += (eigenE[i] * alpha[i]) * eigev[i][0];
            d2[1] += (eigenE[i] * alpha[i]) * eigev[i][1];
            d2[2] += (eigenE[i] * alpha[i]) * eigev[i][2];
        } else {
            d1[0] += (eigenE[i] * alpha[i]) * eigev[i][0];
            d1[1] += (eigenE[i] * alpha[i]) * eigev[i][1];
            d1[2] += (eigenE[i] * alpha[i]) * eigev[i][2];
        }
    }
    for (i = 0; i < 3; i++) {
        if (FABS(d1[i]) < FL_EPS12) d1[i] = (REAL)0.0;
        if (FABS(d2[i]) < FL_EPS12) d2[i] = (REAL)0.0;
    }
    return 1;
}

/* X flux functor (triton_appfis.hpp:608-824) over configX = [0,nx+1) x [1,ny+1):
 * edge (ii, ii+1); rhs*0[ii] = dt*d1/dx, rhs*1[ii+1] = dt*d2/dx (overwrite). */
static void FN(flux_x)(int W, int nx, int ny, REAL dx, REAL dt, REAL hextra, const REAL* h, const REAL* qx,
                       const REAL* qy, const REAL* z, const REAL* sqrth, REAL* r[6])
{
#pragma omp parallel for schedule(dynamic, 1)
    for (int jj = 1; jj < ny + 1; jj++)
        for (int ii = 0; ii < nx + 1; ii++) {
            int id1 = ii, id2 = ii + 1;
            REAL d1[3], d2[3];
            if (!FN(roe)(AT(h, id1, jj), AT(h, id2, jj), AT(qx, id1, jj), AT(qy, id1, jj), AT(z, id1, jj),
                         AT(qx, id2, jj), AT(qy, id2, jj), AT(z, id2, jj), AT(sqrth, id1, jj),
                         AT(sqrth, id2, jj), (REAL)1.0, (REAL)0.0, dx, dt, hextra, d1, d2))
                continue;
            AT(r[0], id1, jj) = dt * d1[0] / dx; /* rhsh0 */
            AT(r[2], id1, jj) = dt * d1[1] / dx; /* rhsqx0 */
            AT(r[4], id1, jj) = dt * d1[2] / dx; /* rhsqy0 */
            AT(r[1], id2, jj) = dt * d2[0] / dx; /* rhsh1 */
            AT(r[3], id2, jj) = dt * d2[1] / dx; /* rhsqx1 */
            AT(r[5], id2, jj) = dt * d2[2] / dx; /* rhsqy1 */
        }
}

/* Y flux functor (triton_appfis.hpp:827-1043) over configY = [1,nx+1) x [1,ny+2):
 * edge (jj, jj-1), normal (0,1); accumulates into rhs*0[jj] and rhs*1[jj-1]. */
static void FN(flux_y)(int W, int nx, int ny, REAL dx, REAL dt, REAL hextra, const REAL* h, const REAL* qx,
                       const REAL* qy, const REAL* z, const REAL* sqrth, REAL* r[6])
{
#pragma omp parallel for schedule(dynamic, 1)
    for (int jj = 1; jj < ny + 2; jj++)
        for (int ii = 1; ii < nx + 1; ii++) {
            int id1 = jj, id2 = jj - 1;
            REAL d1[3], d2[3];
            if (!FN(roe)(AT(h, ii, id1), AT(h, ii, id2), AT(qx, ii, id1), AT(qy, ii, id1), AT(z, ii, id1),
                         AT(qx, ii, id2), AT(qy, ii, id2), AT(z, ii, id2), AT(sqrth, ii, id1),
                         AT(sqrth, ii, id2), (REAL)0.0, (REAL)1.0, dx, dt, hextra, d1, d2))
                continue;
            AT(r[0], ii, id1) = AT(r[0], ii, id1) + dt * d1[0] / dx;
            AT(r[2], ii, id1) = AT(r[2], ii, id1) + dt * d1[1] / dx;
            AT(r[4], ii, id1) = AT(r[4], ii, id1) + dt * d1[2] / dx;
            AT(r[1], ii, id2) = AT(r[1], ii, id2) + dt * d2[0] / dx;
            AT(r[3], ii, id2) = AT(r[3], ii, id2) + dt * d2[1] / dx;
            AT(r[5], ii, id2) = AT(r[5], ii, id2) + dt * d2[2] / dx;
        }
}

/* Update functor with implicit Manning friction (triton_appfis.hpp:1046-1092) over the
 * interior, followed by the config's uniform rainfall h += rain*dt. */
static void FN(update)(int W, int nx, int ny, REAL dt, REAL hextra, REAL rain, REAL* h, REAL* qx, REAL* qy,
                       const REAL* n2, REAL* r[6])
{
#pragma omp parallel for schedule(dynamic, 1)
    for (int j = 1; j <= ny; j++)
        for (int i = 1; i <= nx; i++) {
            REAL hij, qxij, qyij;
            REAL hn = AT(h, i, j);
            hij = hn - AT(r[0], i, j) - AT(r[1], i, j);
            if (hij < (REAL)0.0) hij = (REAL)0.0;
            if (hij < hextra) {
                qxij = (REAL)0.0;
                qyij = (REAL)0.0;
            } else {
                REAL mx = AT(qx, i, j) - AT(r[2], i, j) - AT(r[3], i, j);
                REAL my = AT(qy, i, j) - AT(r[4], i, j) - AT(r[5], i, j);
                REAL modM = SQRT(mx * mx + my * my);
                REAL nn = AT(n2, i, j);
                if (nn > FL_EPS12 && hn >= hextra && modM > FL_EPS12) {
                    REAL tt = dt * FL_G * nn * modM / (hn * hn * CBRT(hn));
                    qxij = (REAL)-0.5 * (mx - mx * SQRT((REAL)1.0 + (REAL)4.0 * tt)) / tt;
                    qyij = (REAL)-0.5 * (my - my * SQRT((REAL)1.0 + (REAL)4.0 * tt)) / tt;
                } else {
                    qxij = mx;
                    qyij = my;
                }
            }
            AT(h, i, j) = hij + rain * dt;
            AT(qx, i, j) = qxij;
            AT(qy, i, j) = qyij;
        }
}

/* Wet/dry momentum fix (triton_appfis.hpp:1095-1112) over the interior. */
static void FN(wetdry)(int W, int nx, int ny, REAL hextra, const REAL* h, REAL* qx, REAL* qy, const REAL* z)
{
#pragma omp parallel for schedule(dynamic, 1)
    for (int j = 1; j <= ny; j++)
        for (int i = 1; i <= nx; i++) {
            REAL hij = AT(h, i, j);
            REAL zij = AT(z, i, j);
            if (hij > hextra) {
                if (((hij + zij < AT(z, i + 1, j)) && (AT(h, i + 1, j) < FL_EPS12))
                    || ((hij + zij < AT(z, i - 1, j)) && (AT(h, i - 1, j) < FL_EPS12)))
                    AT(qx, i, j) = (REAL)0.0;
                if (((hij + zij < AT(z, i, j - 1)) && (AT(h, i, j - 1) < FL_EPS12))
                    || ((hij + zij < AT(z, i, j + 1)) && (AT(h, i, j + 1) < FL_EPS12)))
                    AT(qy, i, j) = (REAL)0.0;
            }
        }
}

/* Point sources (triton_appfis.hpp:358-377 on the host, functor 1115-1141): find the
 * hydrograph interval of simtime starting from *idx_low (kept across steps), interpolate
 * each source's flow linearly in time and add flow*dt/dx^2 to the depth of its cell.
 * src_flow is rows x nsrc, row-major; src_cell are padded linear indices. */
static void FN(sources)(REAL* h, int nsrc, const int* src_cell, int rows, const double* src_time,
                        const double* src_flow, double simtime, double gdt, double dx, int* idx_low)
{
    int lo = *idx_low, hi = rows - 1, found = 0;
    for (int i = lo; i < hi; i++)
        if (src_time[i + 1] > simtime) {
            lo = i;
            hi = i + 1;
            found = 1;
            break;
        }
    if (!found) lo = hi;
    *idx_low = lo;
    for (int s = 0; s < nsrc; s++) {
        double fl = src_flow[(size_t)lo * nsrc + s], fh = src_flow[(size_t)hi * nsrc + s];
        double tl = src_time[lo], th = src_time[hi];
        double flow;
        if (lo == hi) flow = fl;
        else flow = fl + (((fh - fl) * (simtime - tl)) / (th - tl));
        double hij = (double)h[src_cell[s]];
        hij += (flow * gdt) / (dx * dx);
        h[src_cell[s]] = (REAL)hij;
    }
}

/* The step loop of triton_appfis.hpp:316-386 for `steps` steps (fixed count): dynamic
 * dt unless time_increment_fixed (then dt_fallback every step, no print clamp, as the
 * reference skips the CFL block), the config's rain folded into the update pass, then
 * the point sources (nsrc may be 0), the wet/dry fix and the time bookkeeping. */
int FN(orc_flood_run_src)(REAL* h, REAL* qx, REAL* qy, const REAL* z, const REAL* n2, int nx, int ny, double dx_,
                          double cn_, double hextra_, double dt_fallback, double print_interval, double rain_,
                          double start_time, int time_increment_fixed, int nsrc, const int* src_cell, int rows,
                          const double* src_time, const double* src_flow, int steps, double* time_out,
                          double* volume_out, double* dt_out)
{
    const int W = nx + 2, H = ny + 2;
    const size_t n = (size_t)W * (size_t)H;
    REAL* work = (REAL*)calloc(8 * n, sizeof(REAL));
    if (!work) return 1;
    REAL* dtv = work;
    REAL* sqrth = work + n;
    REAL* r[6];
    for (int q = 0; q < 6; q++) r[q] = work + (2 + q) * n;
    const REAL dx = (REAL)dx_, cn = (REAL)cn_, hextra = (REAL)hextra_, rain = (REAL)rain_;
    double simtime = start_time;
    int print_id = 0, idx_low = 0;
    for (int s = 0; s < steps; s++) {
        double gdt = dt_fallback;
        if (!time_increment_fixed) {
            FN(cfl)(W, nx, ny, dx, cn, hextra, qx, qy, h, dtv);
            REAL m = RMAX;
            for (int j = 1; j <= ny; j++)
                for (int i = 1; i <= nx; i++)
                    if (AT(dtv, i, j) < m) m = AT(dtv, i, j);
            gdt = (double)m;
            if (gdt >= (double)RMAX - 1.0) gdt = dt_fallback;
            if (print_interval > 0.0 && simtime + gdt > print_interval * (print_id + 1))
                gdt = print_interval * (print_id + 1) - simtime;
        }
        const REAL dt = (REAL)gdt;
        FN(reset)(W, H, h, sqrth, r);
        FN(flux_x)(W, nx, ny, dx, dt, hextra, h, qx, qy, z, sqrth, r);
        FN(flux_y)(W, nx, ny, dx, dt, hextra, h, qx, qy, z, sqrth, r);
        FN(update)(W, nx, ny, dt, hextra, rain, h, qx, qy, n2, r);
        if (nsrc > 0) FN(sources)(h, nsrc, src_cell, rows, src_time, src_flow, simtime, (double)dt, dx_, &idx_low);
        FN(wetdry)(W, nx, ny, hextra, h, qx, qy, z);
        simtime += (double)dt;
        if (print_interval > 0.0 && simtime >= print_interval * (print_id + 1)) print_id++;
        if (time_out) time_out[s] = simtime;
        if (dt_out) dt_out[s] = (double)dt;
        if (volume_out) {
            long double v = 0.0L;
            for (int j = 1; j <= ny; j++)
                for (int i = 1; i <= nx; i++) v += (long double)AT(h, i, j);
            volume_out[s] = (double)(v * (long double)dx_ * (long double)dx_);
        }
    }
    free(work);
    return 0;
}

/* No point sources, dynamic dt: the config's workload. */
int FN(orc_flood_run)(REAL* h, REAL* qx, REAL* qy, const REAL* z, const REAL* n2, int nx, int ny, double dx_,
                      double cn_, double hextra_, double dt_fallback, double print_interval, double rain_,
                      double start_time, int steps, double* time_out, double* volume_out, double* dt_out)
{
    return FN(orc_flood_run_src)(h, qx, qy, z, n2, nx, ny, dx_, cn_, hextra_, dt_fallback, print_interval, rain_,
                                 start_time, 0, 0, NULL, 0, NULL, NULL, steps, time_out, volume_out, dt_out);
}

#undef AT
#undef FN
#undef FL_G
#undef FL_SQRTG
#undef FL_EPS12
