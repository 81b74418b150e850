/*
 * nyles_oracle.c -- CPU restatement of the Nyles time-step arithmetic.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle the HIP library is
 * checked against.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product path never does.
 *
 * It restates, loop by loop and operation by operation, the Fortran
 * kernels of the reference (pvthinker/Nyles):
 *   core/weno.f90                 weno3, weno5, flux1d
 *   core/fortran_upwind.f90       upwind (WENO branch, and the linear
 *                                 branch with core/interpolate_tracer.f90)
 *   core/fortran_vortex_force.f90 vortex_force_direc / vortex_force_flip
 *                                 (WENO branch, and the linear branch with
 *                                 core/interpolate.f90)
 *   core/fortran_vorticity.f90    vorticity
 *   core/fortran_bernoulli.f90    gradke, gradkeandb, div
 *   core/fortran_kinenergy.f90    kin
 *   core/fortran_dissipation.f90  add_laplacian (viscosity)
 *   core/mgfor/*.f90              the multigrid (setup, solve, vcycle,
 *                                 smoother, residual, restriction,
 *                                 prolongation, norm, halo fill)
 *
 * Arithmetic rules that follow from how the reference is built:
 *  - the f2py kernels are compiled without -fdefault-real-8, so every real
 *    literal in them is a REAL(4) constant widened to double (F32 below);
 *  - tau5 in weno5 is an implicitly typed REAL(4) variable
 *    (core/weno.f90:45), so |beta1-beta3| is rounded to float;
 *  - gfortran on plain x86-64 emits no FMA: this file must be built with
 *    -ffp-contract=off -fno-fast-math;
 *  - mgfor is built with -freal-4-real-8, so its literals are double.
 *
 * Arrays are passed as strided 3-D views: element (a,b,c) of a view with
 * Fortran dims (l,m,n) lives at p[a*s0 + b*s1 + c*s2] (0-based).  The
 * reference physically transposes arrays before each call
 * (core/variables.py:162-182); a strided view is value-identical.
 *
 * The MG arrays keep the mgfor layout x(1-nh:nx+nh, 1-nh:ny+nh, 1:nz) with
 * the z halo inside the array (core/mgfor/mg_setup.f90:15-27,239).  Fortran
 * ALLOCATE leaves memory uninitialised; this restatement zero-fills every
 * MG array at allocation, the value fresh pages carry.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef int64_t i64;

/* REAL(4) literal widened to REAL(8) */
#define F32(x) ((double)(x))

/* ------------------------------------------------------------------ */
/* core/weno.f90:1-22                                                   */
static double weno3(double qm, double q0, double qp)
{
    const double eps = F32(1e-14f);
    double qi1 = (-qm + 3.0 * q0) * 0.5;
    double qi2 = (q0 + qp) * 0.5;
    double beta1 = (q0 - qm) * (q0 - qm);
    double beta2 = (qp - q0) * (qp - q0);
    double tau = fabs(beta2 - beta1);
    double w1 = (1.0 + tau / (beta1 + eps));
    double w2 = (1.0 + tau / (beta2 + eps)) * 2.0;
    return (w1 * qi1 + w2 * qi2) / (w1 + w2);
}

/* core/weno.f90:25-54 */
static double weno5(double qmm, double qm, double q0, double qp, double qpp)
{
    const double eps = F32(1e-16f);
    const double c13 = F32(1.0f / 3.0f), c76 = F32(7.0f / 6.0f);
    const double c116 = F32(11.0f / 6.0f), c16 = F32(1.0f / 6.0f);
    const double c56 = F32(5.0f / 6.0f);
    const double k1 = F32(13.0f / 12.0f), k2 = 0.25;

    double qi1 = c13 * qmm - c76 * qm + c116 * q0;
    double qi2 = -(c16 * qm) + c56 * q0 + c13 * qp;
    double qi3 = c13 * q0 + c56 * qp - c16 * qpp;

    double t1 = qmm - 2.0 * qm + q0, t2 = qmm - 4.0 * qm + 3.0 * q0;
    double beta1 = k1 * (t1 * t1) + k2 * (t2 * t2);
    double t3 = qm - 2.0 * q0 + qp, t4 = qm - qp;
    double beta2 = k1 * (t3 * t3) + k2 * (t4 * t4);
    double t5 = q0 - 2.0 * qp + qpp, t6 = 3.0 * q0 - 4.0 * qp + qpp;
    double beta3 = k1 * (t5 * t5) + k2 * (t6 * t6);

    /* tau5 is implicitly REAL(4) (core/weno.f90:45) */
    double tau5 = (double)(float)fabs(beta1 - beta3);

    double w1 = (1.0 + tau5 / (beta1 + eps));
    double w2 = 6.0 * (1.0 + tau5 / (beta2 + eps));
    double w3 = 3.0 * (1.0 + tau5 / (beta3 + eps));
    return (w1 * qi1 + w2 * qi2 + w3 * qi3) / (w1 + w2 + w3);
}

/* core/weno.f90:106-153.  q and u are 0-based here: q[0] is q(1). */
static void flux1d(const double *u, const double *q, double *flux, int n)
{
#define Q(i) q[(i) - 1]
#define U(i) u[(i) - 1]
#define FL(i) flux[(i) - 1]
    int i;
    if (U(1) > 0) FL(1) = U(1) * Q(1);
    else FL(1) = U(1) * weno3(Q(3), Q(2), Q(1));

    if (U(2) > 0) FL(2) = U(2) * weno3(Q(1), Q(2), Q(3));
    else FL(2) = U(2) * weno5(Q(5), Q(4), Q(3), Q(2), Q(1));

    for (i = 3; i <= n - 3; i++) {
        if (U(i) > 0) FL(i) = U(i) * weno5(Q(i - 2), Q(i - 1), Q(i), Q(i + 1), Q(i + 2));
        else FL(i) = U(i) * weno5(Q(i + 3), Q(i + 2), Q(i + 1), Q(i), Q(i - 1));
    }
    i = n - 2;
    if (U(i) > 0) FL(i) = U(i) * weno5(Q(i - 2), Q(i - 1), Q(i), Q(i + 1), Q(i + 2));
    else FL(i) = U(i) * weno3(Q(i + 2), Q(i + 1), Q(i));
    i = n - 1;
    if (U(i) > 0) FL(i) = U(i) * weno3(Q(i - 1), Q(i), Q(i + 1));
    else FL(i) = U(i) * Q(i + 1);
    FL(n) = 0.0;
#undef Q
#undef U
#undef FL
}

/* Exported scalar entry points used by the fixture tests. */
double orc_weno3(double qm, double q0, double qp) { return weno3(qm, q0, qp); }
double orc_weno5(double a, double b, double c, double d, double e) { return weno5(a, b, c, d, e); }
void orc_flux1d(const double *u, const double *q, double *flux, int n)
{
    /* the reference reads q(5) / q(i+3) only when n >= 6; keep callers honest */
    if (n < 6) { for (int i = 0; i < n; i++) flux[i] = NAN; return; }
    flux1d(u, q, flux, n);
}

#define AT(p, a, b, c) (p)[(i64)(a) * s0 + (i64)(b) * s1 + (i64)(c) * s2]

/* core/fortran_upwind.f90:3-87 (WENO branch :221-235).
 * trac,u,dtrac: dims (l,m,n); lines run along the third index. */
void orc_upwind(const double *trac, const double *u, double *dtrac,
                int l, int m, int n, i64 s0, i64 s1, i64 s2)
{
    #pragma omp parallel
    {
        double *buf = (double *)malloc(sizeof(double) * 3 * (size_t)n);
        double *up = buf, *phi = buf + n, *flux = buf + 2 * n;
        #pragma omp for collapse(2) schedule(static)
        for (int k = 0; k < l; k++)
            for (int j = 0; j < m; j++) {
                for (int i = 0; i < n; i++) {
                    up[i] = AT(u, k, j, i);
                    phi[i] = AT(trac, k, j, i);
                }
                flux1d(up, phi, flux, n);
                AT(dtrac, k, j, 0) = AT(dtrac, k, j, 0) - flux[0];
                for (int i = 1; i < n; i++)
                    AT(dtrac, k, j, i) = AT(dtrac, k, j, i) + flux[i - 1] - flux[i];
            }
        free(buf);
    }
}

/* Linear upwind interpolation, odd orders 1, 3, 5, on a line q[0..n-1]
 * (0-based).  Constants are the REAL(4) literals of the Fortran widened to
 * double; the sums run left to right as written there.
 * Tracer flavour, core/interpolate_tracer.f90:3-70: qp[i], qm[i] for
 * i = 0..n-1, the first-order copy at both ends, third order next to them,
 * fifth order in between. */
static void interp_tracer(const double *q, double *qp, double *qm, int order, int n)
{
    const double c1 = -(double)(1.0f / 6.0f), c2 = (double)(5.0f / 6.0f), c3 = (double)(2.0f / 6.0f);
    const double b1 = (double)(2.0f / 60.0f), b2 = -(double)(13.0f / 60.0f);
    const double b3 = (double)(47.0f / 60.0f), b4 = (double)(27.0f / 60.0f);
    const double b5 = -(double)(3.0f / 60.0f);
    for (int i = 0; i < n; i++) { qp[i] = q[i]; qm[i] = q[i]; }
    if (order == 1) return;
    const int lo = order == 5 ? 2 : 1, hi = order == 5 ? n - 3 : n - 2;
    if (order == 5) {
        for (int i = lo; i <= hi; i++) {
            qp[i] = b1 * q[i - 2] + b2 * q[i - 1] + b3 * q[i] + b4 * q[i + 1] + b5 * q[i + 2];
            qm[i] = b5 * q[i - 2] + b4 * q[i - 1] + b3 * q[i] + b2 * q[i + 1] + b1 * q[i + 2];
        }
        const int e[2] = {1, n - 2};
        for (int t = 0; t < 2; t++) {
            const int i = e[t];
            qp[i] = c1 * q[i - 1] + c2 * q[i] + c3 * q[i + 1];
            qm[i] = c3 * q[i - 1] + c2 * q[i] + c1 * q[i + 1];
        }
    } else {
        for (int i = lo; i <= hi; i++) {
            qp[i] = c1 * q[i - 1] + c2 * q[i] + c3 * q[i + 1];
            qm[i] = c3 * q[i - 1] + c2 * q[i] + c1 * q[i + 1];
        }
    }
}

/* Vortex-force flavour, core/interpolate.f90:3-79 (odd orders): qp is
 * indexed 0..n-1 with qp[0] = 0 standing for Fortran's qp(0), so Fortran
 * qp(i) is qp[i] here and Fortran qm(i) is qm[i-1]; that is qp[1..] and
 * qm[0..] are the 1-based values shifted by one.  To keep one convention
 * this takes 1-based arrays: qp[0..n-1] = Fortran qp(0..n-1) and
 * qm[1..n] = Fortran qm(1..n) (qm[0] unused). */
static void interp_vf(const double *v /* 1-based: v[1..n] */, double *qp, double *qm,
                      int order, int n)
{
    const double c1 = -(double)(1.0f / 6.0f), c2 = (double)(5.0f / 6.0f), c3 = (double)(2.0f / 6.0f);
    const double b1 = (double)(2.0f / 60.0f), b2 = -(double)(13.0f / 60.0f);
    const double b3 = (double)(47.0f / 60.0f), b4 = (double)(27.0f / 60.0f);
    const double b5 = -(double)(3.0f / 60.0f);
    qp[0] = 0.0;
    if (order == 5) {
        qp[1] = v[1]; qm[1] = v[1];
        qp[2] = c1 * v[1] + c2 * v[2] + c3 * v[3];
        qm[2] = c3 * v[1] + c2 * v[2] + c1 * v[3];
        for (int i = 3; i <= n - 3; i++) {
            qp[i] = b1 * v[i - 2] + b2 * v[i - 1] + b3 * v[i] + b4 * v[i + 1] + b5 * v[i + 2];
            qm[i] = b5 * v[i - 2] + b4 * v[i - 1] + b3 * v[i] + b2 * v[i + 1] + b1 * v[i + 2];
        }
        int i = n - 2;
        qp[i] = c1 * v[i - 1] + c2 * v[i] + c3 * v[i + 1];
        qm[i] = c3 * v[i - 1] + c2 * v[i] + c1 * v[i + 1];
        i = n - 1;
        qp[i] = v[i]; qm[i] = v[i];
        qm[n] = v[n];
    } else if (order == 3) {
        qp[1] = v[1]; qm[1] = v[1];
        for (int i = 2; i <= n - 2; i++) {
            qp[i] = c1 * v[i - 1] + c2 * v[i] + c3 * v[i + 1];
            qm[i] = c3 * v[i - 1] + c2 * v[i] + c1 * v[i + 1];
        }
        qp[n - 1] = v[n - 1]; qm[n - 1] = v[n - 1];
        qm[n] = v[n];
    } else {
        for (int i = 1; i <= n - 1; i++) { qp[i] = v[i]; qm[i] = v[i]; }
        qm[n] = v[n];
    }
}

void orc_interp_tracer(const double *q, double *qp, double *qm, int order, int n)
{
    interp_tracer(q, qp, qm, order, n);
}

void orc_interp_vf(const double *v, double *qp, double *qm, int order, int n)
{
    interp_vf(v, qp, qm, order, n);
}

/* core/fortran_upwind.f90:34-67, the linear branch, odd orders: per line
 * fx = up(i)*qp(i) + um(i)*qm(i+1) for the faces i = 1..n-1, with
 * up/um = 0.5*(u -+ |u|); dtrac(i) = dtrac(i) + fxm - fx, and fx = 0 past
 * the last face. */
void orc_upwind_linear(const double *trac, const double *u, double *dtrac, int order,
                       int l, int m, int n, i64 s0, i64 s1, i64 s2)
{
    #pragma omp parallel
    {
        double *buf = (double *)malloc(sizeof(double) * 5 * (size_t)n);
        double *up = buf, *um = buf + n, *phi = buf + 2 * n, *qp = buf + 3 * n, *qm = buf + 4 * n;
        #pragma omp for collapse(2) schedule(static)
        for (int k = 0; k < l; k++)
            for (int j = 0; j < m; j++) {
                for (int i = 0; i < n; i++) {
                    const double ui = AT(u, k, j, i), UU = fabs(ui);
                    up[i] = 0.5 * (ui + UU);
                    um[i] = 0.5 * (ui - UU);
                    phi[i] = AT(trac, k, j, i);
                }
                interp_tracer(phi, qp, qm, order, n);
                double fxm = 0.0;
                for (int i = 0; i < n - 1; i++) {
                    const double fx = up[i] * qp[i] + um[i] * qm[i + 1];
                    AT(dtrac, k, j, i) = AT(dtrac, k, j, i) + fxm - fx;
                    fxm = fx;
                }
                AT(dtrac, k, j, n - 1) = AT(dtrac, k, j, n - 1) + fxm - 0.0;
            }
        free(buf);
    }
}

/* core/fortran_vortex_force.f90:43-65 (direc) and :121-144 (flip), the
 * linear branch, odd orders: vU(k) = vort(k) (not shifted), U_interp as in
 * the WENO branch, up/um its split; direc res(k) = res(k) - qp(k-1)*up(k)
 * - qm(k)*um(k), flip res(i) = res(i) + qp(i-1)*up(i) + qm(i)*um(i). */
void orc_vortex_force_direc_linear(const double *U, const double *vort, double *res, int order,
                                   int m, int n, int l, i64 s0, i64 s1, i64 s2)
{
    #pragma omp parallel
    {
        double *buf = (double *)malloc(sizeof(double) * 5 * ((size_t)l + 1));
        double *up = buf, *um = buf + (l + 1), *v = buf + 2 * (l + 1);
        double *qp = buf + 3 * (l + 1), *qm = buf + 4 * (l + 1);
        #pragma omp for collapse(2) schedule(static)
        for (int j = 0; j < m; j++)
            for (int i = 0; i < n - 1; i++) {
                double UU0 = 0.0;
                for (int k = 1; k <= l; k++) {
                    const double UU1 = 0.5 * (AT(U, j, i, k - 1) + AT(U, j, i + 1, k - 1));
                    v[k] = AT(vort, j, i, k - 1);
                    const double Ui = 0.5 * (UU0 + UU1), UU = fabs(Ui);
                    up[k] = 0.5 * (Ui + UU);
                    um[k] = 0.5 * (Ui - UU);
                    UU0 = UU1;
                }
                interp_vf(v, qp, qm, order, l);
                for (int k = 1; k <= l; k++)
                    AT(res, j, i, k - 1) = AT(res, j, i, k - 1) - qp[k - 1] * up[k] - qm[k] * um[k];
            }
        free(buf);
    }
}

void orc_vortex_force_flip_linear(const double *U, const double *vort, double *res, int order,
                                  int m, int n, int l, i64 s0, i64 s1, i64 s2)
{
    #pragma omp parallel
    {
        double *buf = (double *)malloc(sizeof(double) * 5 * ((size_t)n + 1));
        double *up = buf, *um = buf + (n + 1), *v = buf + 2 * (n + 1);
        double *qp = buf + 3 * (n + 1), *qm = buf + 4 * (n + 1);
        #pragma omp for collapse(2) schedule(static)
        for (int j = 0; j < m; j++)
            for (int k = 0; k < l - 1; k++) {
                double UU0 = 0.0;
                for (int i = 1; i <= n; i++) {
                    const double UU1 = 0.5 * (AT(U, j, i - 1, k) + AT(U, j, i - 1, k + 1));
                    v[i] = AT(vort, j, i - 1, k);
                    const double Ui = 0.5 * (UU0 + UU1), UU = fabs(Ui);
                    up[i] = 0.5 * (Ui + UU);
                    um[i] = 0.5 * (Ui - UU);
                    UU0 = UU1;
                }
                interp_vf(v, qp, qm, order, n);
                for (int i = 1; i <= n; i++)
                    AT(res, j, i - 1, k) = AT(res, j, i - 1, k) + qp[i - 1] * up[i] + qm[i] * um[i];
            }
        free(buf);
    }
}

/* core/fortran_vortex_force.f90:10-85 (WENO branch :66-81).
 * U,vort,res: dims (m,n,l) indexed (j,i,k); flux along k, U averaged in i. */
void orc_vortex_force_direc(const double *U, const double *vort, double *res,
                            int m, int n, int l, i64 s0, i64 s1, i64 s2)
{
    #pragma omp parallel
    {
        double *buf = (double *)malloc(sizeof(double) * 3 * (size_t)l);
        double *u1d = buf, *q = buf + l, *flux = buf + 2 * l;
        #pragma omp for collapse(2) schedule(static)
        for (int j = 0; j < m; j++)
            for (int i = 0; i < n - 1; i++) {
                double UU0 = 0.0;
                for (int k = 0; k < l; k++) {
                    double UU1 = 0.5 * (AT(U, j, i, k) + AT(U, j, i + 1, k));
                    u1d[k] = 0.5 * (UU0 + UU1);
                    UU0 = UU1;
                }
                q[0] = 0.0;
                for (int k = 1; k < l; k++) q[k] = AT(vort, j, i, k - 1);
                flux1d(u1d, q, flux, l);
                for (int k = 0; k < l; k++) AT(res, j, i, k) = AT(res, j, i, k) - flux[k];
            }
        free(buf);
    }
}

/* core/fortran_vortex_force.f90:87-165 (WENO branch :145-159).
 * flux along i (second index), U averaged in k (third index). */
void orc_vortex_force_flip(const double *U, const double *vort, double *res,
                           int m, int n, int l, i64 s0, i64 s1, i64 s2)
{
    #pragma omp parallel
    {
        double *buf = (double *)malloc(sizeof(double) * 3 * (size_t)n);
        double *u1d = buf, *q = buf + n, *flux = buf + 2 * n;
        #pragma omp for collapse(2) schedule(static)
        for (int j = 0; j < m; j++)
            for (int k = 0; k < l - 1; k++) {
                double UU0 = 0.0;
                for (int i = 0; i < n; i++) {
                    double UU1 = 0.5 * (AT(U, j, i, k) + AT(U, j, i, k + 1));
                    u1d[i] = 0.5 * (UU0 + UU1);
                    UU0 = UU1;
                }
                q[0] = 0.0;
                for (int i = 1; i < n; i++) q[i] = AT(vort, j, i - 1, k);
                flux1d(u1d, q, flux, n);
                for (int i = 0; i < n; i++) AT(res, j, i, k) = AT(res, j, i, k) + flux[i];
            }
        free(buf);
    }
}

/* core/fortran_vorticity.f90:2-28 */
void orc_vorticity(const double *ui, const double *uj, double *wk,
                   int l, int m, int n, i64 s0, i64 s1, i64 s2)
{
    #pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < l; k++)
        for (int j = 0; j < m - 1; j++) {
            for (int i = 0; i < n - 1; i++)
                AT(wk, k, j, i) = AT(uj, k, j, i + 1) - AT(uj, k, j, i)
                                - AT(ui, k, j + 1, i) + AT(ui, k, j, i);
            AT(wk, k, j, n - 1) = 0.0;
        }
}

/* core/fortran_bernoulli.f90:2-26 */
void orc_gradke(const double *ke, double *du, int l, int m, int n, i64 s0, i64 s1, i64 s2)
{
    #pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < l; k++)
        for (int j = 0; j < m; j++)
            for (int i = 0; i < n - 1; i++)
                AT(du, k, j, i) = AT(du, k, j, i) - (AT(ke, k, j, i + 1) - AT(ke, k, j, i));
}

/* core/fortran_bernoulli.f90:29-58 */
void orc_gradkeandb(const double *ke, const double *b, double *du, double dz,
                    int l, int m, int n, i64 s0, i64 s1, i64 s2)
{
    const double cff = 0.5 * dz;
    #pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < l; k++)
        for (int j = 0; j < m; j++)
            for (int i = 0; i < n - 1; i++)
                AT(du, k, j, i) = AT(du, k, j, i) - (AT(ke, k, j, i + 1) - AT(ke, k, j, i))
                                + cff * (AT(b, k, j, i + 1) + AT(b, k, j, i));
}

/* core/fortran_bernoulli.f90:61-97 */
void orc_div(double *d, const double *u, int iflag, int l, int m, int n, i64 s0, i64 s1, i64 s2)
{
    #pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < l; k++)
        for (int j = 0; j < m; j++) {
            if (iflag > 0) {
                AT(d, k, j, 0) = AT(d, k, j, 0) + AT(u, k, j, 0);
                for (int i = 1; i < n; i++)
                    AT(d, k, j, i) = AT(d, k, j, i) + (AT(u, k, j, i) - AT(u, k, j, i - 1));
            } else {
                AT(d, k, j, 0) = AT(u, k, j, 0);
                for (int i = 1; i < n; i++)
                    AT(d, k, j, i) = (AT(u, k, j, i) - AT(u, k, j, i - 1));
            }
        }
}

/* core/fortran_kinenergy.f90:3-56: ke += (cff2*0.5)*(u(i)**2+u(i-1)**2), i>=2 */
void orc_kin(const double *u, double *ke, double ds2, int l, int m, int n, i64 s0, i64 s1, i64 s2)
{
    const double cff2 = 0.5 * ds2;
    #pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < l; k++)
        for (int j = 0; j < m; j++)
            for (int i = 1; i < n; i++)
                AT(ke, k, j, i) = AT(ke, k, j, i)
                    + cff2 * 0.5 * (AT(u, k, j, i) * AT(u, k, j, i) + AT(u, k, j, i - 1) * AT(u, k, j, i - 1));
}
/* core/fortran_dissipation.f90:2-34 add_laplacian: dphi += coef*delta[delta[phi]]
 * along the third index, with no flux through either end of a line (fxm = 0
 * at the first cell, fx = 0 at the last). */
void orc_add_laplacian(const double *phi, double *dphi, double coef,
                       int l, int m, int n, i64 s0, i64 s1, i64 s2)
{
    #pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < l; k++)
        for (int j = 0; j < m; j++) {
            double fxm = 0.0, fx;
            for (int i = 0; i < n - 1; i++) {
                fx = AT(phi, k, j, i + 1) - AT(phi, k, j, i);
                AT(dphi, k, j, i) = AT(dphi, k, j, i) + coef * (fx - fxm);
                fxm = fx;
            }
            fx = 0.0;
            AT(dphi, k, j, n - 1) = AT(dphi, k, j, n - 1) + coef * (fx - fxm);
        }
}
#undef AT

/* ================================================================== */
/*  Multigrid: core/mgfor                                              */
/* ================================================================== */

enum { T_CLOSED = 1, T_XPERIO, T_YPERIO, T_ZPERIO, T_XYPERIO, T_XYZPERIO };

typedef struct {
    int nx, ny, nz;          /* nz = array z extent (interior + 2 nh) */
    int nh;
    i64 rs, ps;              /* row and plane strides */
    i64 n;                   /* total elements */
    double *x, *b, *r, *y, *diag, *idiag, *Rcoef, *Pcoef;
    int *msk;
} orc_level;

typedef struct {
    int nlevels;
    int topology;
    int xperio, yperio, zperio;
    int nh;
    int maxite;
    double tol, omega;
    orc_level lev[50];
    /* statistics of the last solve */
    int nite;
    double res, normb;
    int total_cycles;
    double *tmp;             /* scratch for fills */
} orc_mg;

/* element (i,j,k) in Fortran indices, i in 1-nh..nx+nh, j alike, k in 1..nz */
#define MI(L, i, j, k) ((i64)((k) - 1) * (L)->ps + (i64)((j) + (L)->nh - 1) * (L)->rs + ((i) + (L)->nh - 1))

/* core/mgfor/mod_halo.f90:235-262 (single rank: every neighbour is myself).
 * Fortran array assignment evaluates the right-hand side first: copy
 * through a temporary so overlapping sections (n < nh) behave the same. */
static void fill_d(orc_mg *mg, orc_level *L, double *a)
{
    const int nx = L->nx, ny = L->ny, nz = L->nz, nh = L->nh;
    double *t = mg->tmp;
    int i, j, k;
#define CP_SECTION(i0, i1, j0, j1, k0, k1, si0, sj0, sk0)                     \
    do {                                                                       \
        i64 c = 0;                                                             \
        for (k = 0; k <= (k1) - (k0); k++)                                     \
            for (j = 0; j <= (j1) - (j0); j++)                                 \
                for (i = 0; i <= (i1) - (i0); i++)                             \
                    t[c++] = a[MI(L, (si0) + i, (sj0) + j, (sk0) + k)];        \
        c = 0;                                                                 \
        for (k = 0; k <= (k1) - (k0); k++)                                     \
            for (j = 0; j <= (j1) - (j0); j++)                                 \
                for (i = 0; i <= (i1) - (i0); i++)                             \
                    a[MI(L, (i0) + i, (j0) + j, (k0) + k)] = t[c++];           \
    } while (0)

    if (mg->xperio) {
        CP_SECTION(nx + 1, nx + nh, 1, ny, 1, nz, 1, 1, 1);
        CP_SECTION(1 - nh, 0, 1, ny, 1, nz, nx - nh + 1, 1, 1);
    }
    if (mg->yperio) {
        CP_SECTION(1, nx, ny + 1, ny + nh, 1, nz, 1, 1, 1);
        CP_SECTION(1, nx, 1 - nh, 0, 1, nz, 1, ny - nh + 1, 1);
    }
    if (mg->xperio && mg->yperio) {
        CP_SECTION(nx + 1, nx + nh, ny + 1, ny + nh, 1, nz, 1, 1, 1);
        CP_SECTION(1 - nh, 0, ny + 1, ny + nh, 1, nz, nx - nh + 1, 1, 1);
        CP_SECTION(nx + 1, nx + nh, 1 - nh, 0, 1, nz, 1, ny - nh + 1, 1);
        CP_SECTION(1 - nh, 0, 1 - nh, 0, 1, nz, nx - nh + 1, ny - nh + 1, 1);
    }
    if (mg->zperio) {
        CP_SECTION(1 - nh, nx + nh, 1 - nh, ny + nh, 1, nh, 1 - nh, 1 - nh, nz - 2 * nh + 1);
        CP_SECTION(1 - nh, nx + nh, 1 - nh, ny + nh, nz - nh + 1, nz, 1 - nh, 1 - nh, nh + 1);
    }
#undef CP_SECTION
}

static void fill_i(orc_mg *mg, orc_level *L, int *m)
{
    /* msk fills go through the REAL y array in the reference
       (core/mgfor/mg_setup.f90:213-223); integers 0/1 are exact. */
    for (i64 q = 0; q < L->n; q++) L->y[q] = (double)m[q];
    fill_d(mg, L, L->y);
    for (i64 q = 0; q < L->n; q++) m[q] = (int)L->y[q];
}

/* core/mgfor/basicoperators.f90:32-60 */
static void frestrict(const orc_level *F, const double *xf, const orc_level *C, double *xc, const double *coef)
{
    const int nh = C->nh;
    #pragma omp parallel for schedule(static)
    for (int kc = 1 + nh; kc <= C->nz - nh; kc++) {
        int k = 1 + nh + (kc - 1 - nh) * 2;
        for (int jc = 1; jc <= C->ny; jc++) {
            int j = 1 + (jc - 1) * 2;
            for (int ic = 1; ic <= C->nx; ic++) {
                int i = 1 + (ic - 1) * 2;
                xc[MI(C, ic, jc, kc)] = coef[MI(C, ic, jc, kc)] * (
                    xf[MI(F, i, j, k)] + xf[MI(F, i + 1, j, k)]
                  + xf[MI(F, i, j + 1, k)] + xf[MI(F, i + 1, j + 1, k)]
                  + xf[MI(F, i, j, k + 1)] + xf[MI(F, i + 1, j, k + 1)]
                  + xf[MI(F, i, j + 1, k + 1)] + xf[MI(F, i + 1, j + 1, k + 1)]);
            }
        }
    }
}

/* core/mgfor/basicoperators.f90:173-231 */
static void fprolongation(const orc_level *F, double *xf, const orc_level *C, const double *xc, const double *coef)
{
    const int nh = F->nh;
    #pragma omp parallel for schedule(static)
    for (int k = 1 + nh; k <= F->nz - nh; k += 2) {
        int kc = nh + (k + 1 - nh) / 2;
        for (int j = 1; j <= F->ny - 1; j += 2) {
            int jc = (j + 1) / 2;
            for (int i = 1; i <= F->nx - 1; i += 2) {
                int ic = (i + 1) / 2;
                double a, b, c;
#define XC(ii, jj, kk) xc[MI(C, ii, jj, kk)]
#define XF(ii, jj, kk) xf[MI(F, ii, jj, kk)]
#define CF(ii, jj, kk) coef[MI(F, ii, jj, kk)]
                a = 9.0 * XC(ic, jc, kc - 1) + 3.0 * XC(ic - 1, jc, kc - 1) + 3.0 * XC(ic, jc - 1, kc - 1) + XC(ic - 1, jc - 1, kc - 1);
                b = 9.0 * XC(ic, jc, kc) + 3.0 * XC(ic - 1, jc, kc) + 3.0 * XC(ic, jc - 1, kc) + XC(ic - 1, jc - 1, kc);
                c = 9.0 * XC(ic, jc, kc + 1) + 3.0 * XC(ic - 1, jc, kc + 1) + 3.0 * XC(ic, jc - 1, kc + 1) + XC(ic - 1, jc - 1, kc + 1);
                XF(i, j, k) = XF(i, j, k) + CF(i, j, k) * (3.0 * b + a);
                XF(i, j, k + 1) = XF(i, j, k + 1) + CF(i, j, k + 1) * (3.0 * b + c);

                a = 9.0 * XC(ic, jc, kc - 1) + 3.0 * XC(ic + 1, jc, kc - 1) + 3.0 * XC(ic, jc - 1, kc - 1) + XC(ic + 1, jc - 1, kc - 1);
                b = 9.0 * XC(ic, jc, kc) + 3.0 * XC(ic + 1, jc, kc) + 3.0 * XC(ic, jc - 1, kc) + XC(ic + 1, jc - 1, kc);
                c = 9.0 * XC(ic, jc, kc + 1) + 3.0 * XC(ic + 1, jc, kc + 1) + 3.0 * XC(ic, jc - 1, kc + 1) + XC(ic + 1, jc - 1, kc + 1);
                XF(i + 1, j, k) = XF(i + 1, j, k) + CF(i + 1, j, k) * (3.0 * b + a);
                XF(i + 1, j, k + 1) = XF(i + 1, j, k + 1) + CF(i + 1, j, k + 1) * (3.0 * b + c);

                a = 9.0 * XC(ic, jc, kc - 1) + 3.0 * XC(ic - 1, jc, kc - 1) + 3.0 * XC(ic, jc + 1, kc - 1) + XC(ic - 1, jc + 1, kc - 1);
                b = 9.0 * XC(ic, jc, kc) + 3.0 * XC(ic - 1, jc, kc) + 3.0 * XC(ic, jc + 1, kc) + XC(ic - 1, jc + 1, kc);
                c = 9.0 * XC(ic, jc, kc + 1) + 3.0 * XC(ic - 1, jc, kc + 1) + 3.0 * XC(ic, jc + 1, kc + 1) + XC(ic - 1, jc + 1, kc + 1);
                XF(i, j + 1, k) = XF(i, j + 1, k) + CF(i, j + 1, k) * (3.0 * b + a);
                XF(i, j + 1, k + 1) = XF(i, j + 1, k + 1) + CF(i, j + 1, k + 1) * (3.0 * b + c);

                a = 9.0 * XC(ic, jc, kc - 1) + 3.0 * XC(ic + 1, jc, kc - 1) + 3.0 * XC(ic, jc + 1, kc - 1) + XC(ic + 1, jc + 1, kc - 1);
                b = 9.0 * XC(ic, jc, kc) + 3.0 * XC(ic + 1, jc, kc) + 3.0 * XC(ic, jc + 1, kc) + XC(ic + 1, jc + 1, kc);
                c = 9.0 * XC(ic, jc, kc + 1) + 3.0 * XC(ic + 1, jc, kc + 1) + 3.0 * XC(ic, jc + 1, kc + 1) + XC(ic + 1, jc + 1, kc + 1);
                XF(i + 1, j + 1, k) = XF(i + 1, j + 1, k) + CF(i + 1, j + 1, k) * (3.0 * b + a);
                XF(i + 1, j + 1, k + 1) = XF(i + 1, j + 1, k + 1) + CF(i + 1, j + 1, k + 1) * (3.0 * b + c);
#undef XC
#undef XF
#undef CF
            }
        }
    }
}

/* core/mgfor/basicoperators.f90:300-323 */
static void fresidual(orc_level *L, const double *x, const double *b, double *r, const int *msk, const double *diag)
{
    const int nh = L->nh;
    #pragma omp parallel for schedule(static)
    for (int k = 1 + nh; k <= L->nz - nh; k++)
        for (int j = 1; j <= L->ny; j++)
            for (int i = 1; i <= L->nx; i++) {
                i64 q = MI(L, i, j, k);
                r[q] = (double)msk[q] * (b[q] + diag[q] * x[q]
                        - (x[MI(L, i - 1, j, k)] + x[MI(L, i + 1, j, k)]
                         + x[MI(L, i, j - 1, k)] + x[MI(L, i, j + 1, k)]
                         + x[MI(L, i, j, k - 1)] + x[MI(L, i, j, k + 1)]));
            }
}

/* core/mgfor/basicoperators.f90:363-400 */
static void fsmoother(orc_level *L, double *x, const double *b, double *y, const double *idiag, double omega)
{
    const int nh = L->nh;
    const double cff1 = 1.0 - omega;
    #pragma omp parallel for schedule(static)
    for (int k = nh; k <= L->nz + 1 - nh; k++)
        for (int j = 0; j <= L->ny + 1; j++)
            for (int i = 0; i <= L->nx + 1; i++) {
                i64 q = MI(L, i, j, k);
                y[q] = (cff1 * x[q] + omega * ((x[MI(L, i - 1, j, k)] + x[MI(L, i + 1, j, k)]
                        + x[MI(L, i, j - 1, k)] + x[MI(L, i, j + 1, k)]
                        + x[MI(L, i, j, k - 1)] + x[MI(L, i, j, k + 1)]) - b[q]) * idiag[q]);
            }
    #pragma omp parallel for schedule(static)
    for (int k = nh + 1; k <= L->nz - nh; k++)
        for (int j = 1; j <= L->ny; j++)
            for (int i = 1; i <= L->nx; i++) {
                i64 q = MI(L, i, j, k);
                x[q] = (cff1 * y[q] + omega * ((y[MI(L, i - 1, j, k)] + y[MI(L, i + 1, j, k)]
                        + y[MI(L, i, j - 1, k)] + y[MI(L, i, j + 1, k)]
                        + y[MI(L, i, j, k - 1)] + y[MI(L, i, j, k + 1)]) - b[q]) * idiag[q]);
            }
}

/* core/mgfor/basicoperators.f90:422-440: sequential sum in loop order */
static double fnorm(orc_level *L, const int *msk, const double *x)
{
    const int nh = L->nh;
    double sum = 0.0;
    for (int k = 1 + nh; k <= L->nz - nh; k++)
        for (int j = 1; j <= L->ny; j++)
            for (int i = 1; i <= L->nx; i++) {
                i64 q = MI(L, i, j, k);
                sum = sum + (double)msk[q] * (x[q] * x[q]);
            }
    return sum;
}

/* core/mgfor/operators.f90:127-149 */
static void residual(orc_mg *mg, int lev)
{
    orc_level *L = &mg->lev[lev];
    fresidual(L, L->x, L->b, L->r, L->msk, L->diag);
    fill_d(mg, L, L->r);
}

/* core/mgfor/operators.f90:151-170: one fsmoother3d call whatever nite is */
static void smooth(orc_mg *mg, int lev)
{
    orc_level *L = &mg->lev[lev];
    fsmoother(L, L->x, L->b, L->y, L->idiag, mg->omega);
    fill_d(mg, L, L->x);
}

/* core/mgfor/operators.f90:172-213 (no gluing on one rank) */
static void restriction(orc_mg *mg, int lev, int which_rb)
{
    orc_level *F = &mg->lev[lev], *C = &mg->lev[lev + 1];
    const double *fine = which_rb ? F->r : F->b;
    frestrict(F, fine, C, C->b, C->Rcoef);
    memset(C->x, 0, sizeof(double) * (size_t)C->n);
    fill_d(mg, C, C->b);
}

/* core/mgfor/operators.f90:215-244 */
static void prolongation(orc_mg *mg, int lev)
{
    orc_level *F = &mg->lev[lev], *C = &mg->lev[lev + 1];
    fprolongation(F, F->x, C, C->x, F->Pcoef);
    fill_d(mg, F, F->x);
}

/* core/mgfor/operators.f90:98-125 */
static double norm_of(orc_mg *mg, int lev, int which_r)
{
    orc_level *L = &mg->lev[lev];
    return fnorm(L, L->msk, which_r ? L->r : L->b);
}

/* core/mgfor/operators.f90:81-96 */
static double normresidual(orc_mg *mg, double normb)
{
    if (normb > 0.0) {
        residual(mg, 0);
        return norm_of(mg, 0, 1) / normb;
    }
    return 0.0;
}

/* core/mgfor/solvers.f90:35-55 */
static void vcycle(orc_mg *mg)
{
    int lev1 = mg->nlevels - 1;          /* 0-based coarsest */
    for (int lev = 0; lev < lev1; lev++) {
        smooth(mg, lev);
        residual(mg, lev);
        restriction(mg, lev, 1);
    }
    smooth(mg, lev1);
    for (int lev = lev1 - 1; lev >= 0; lev--) {
        prolongation(mg, lev);
        smooth(mg, lev);
    }
}

/* core/mgfor/solvers.f90:8-33 */
int orc_mg_solve(orc_mg *mg)
{
    int nite = 0;
    double normb = norm_of(mg, 0, 0);
    double res = normresidual(mg, normb);
    for (;;) {
        if (res < mg->tol) break;
        vcycle(mg);
        nite = nite + 1;
        if (nite >= mg->maxite) break;
        res = normresidual(mg, normb);
    }
    mg->nite = nite;
    mg->res = res;
    mg->normb = normb;
    mg->total_cycles += nite;
    return nite;
}

void orc_mg_vcycle(orc_mg *mg) { vcycle(mg); }
void orc_mg_smooth(orc_mg *mg, int lev) { smooth(mg, lev); }
void orc_mg_residual(orc_mg *mg, int lev) { residual(mg, lev); }
void orc_mg_restrict(orc_mg *mg, int lev) { restriction(mg, lev, 1); }
void orc_mg_prolong(orc_mg *mg, int lev) { prolongation(mg, lev); }
double orc_mg_norm(orc_mg *mg, int lev, int which_r) { return norm_of(mg, lev, which_r); }
void orc_mg_fill(orc_mg *mg, int lev, double *a) { fill_d(mg, &mg->lev[lev], a); }

/* --- setup: core/mgfor/mg_setup.f90 and core/mgfor/operators.f90:246-505 --- */

static void alloc_level(orc_level *L, int nx, int ny, int nz, int nh)
{
    L->nx = nx; L->ny = ny; L->nz = nz; L->nh = nh;
    L->rs = nx + 2 * nh;
    L->ps = (i64)(nx + 2 * nh) * (ny + 2 * nh);
    L->n = L->ps * nz;
    L->x = (double *)calloc(L->n, sizeof(double));
    L->b = (double *)calloc(L->n, sizeof(double));
    L->r = (double *)calloc(L->n, sizeof(double));
    L->y = (double *)calloc(L->n, sizeof(double));
    L->diag = (double *)calloc(L->n, sizeof(double));
    L->idiag = (double *)calloc(L->n, sizeof(double));
    L->Rcoef = (double *)calloc(L->n, sizeof(double));
    L->Pcoef = (double *)calloc(L->n, sizeof(double));
    L->msk = (int *)calloc(L->n, sizeof(int));
}

/* core/mgfor/operators.f90:246-297 */
static void apply_default_msk(orc_mg *mg, int lev)
{
    orc_level *L = &mg->lev[lev];
    for (i64 q = 0; q < L->n; q++) L->y[q] = 1.0;
    for (int k = 1; k <= L->nz; k++)
        for (int j = 1 - L->nh; j <= L->ny + L->nh; j++)
            for (int i = 1 - L->nh; i <= L->nx + L->nh; i++) {
                int zero = 0;
                if (!mg->xperio && (i <= 0 || i >= L->nx + 1)) zero = 1;
                if (!mg->yperio && (j <= 0 || j >= L->ny + 1)) zero = 1;
                if (zero) L->y[MI(L, i, j, k)] = 0.0;
            }
    for (i64 q = 0; q < L->n; q++) L->msk[q] = (int)((double)L->msk[q] * L->y[q]);
    memset(L->y, 0, sizeof(double) * (size_t)L->n);
}

/* core/mgfor/operators.f90:299-335 */
static void compute_msk(orc_mg *mg, int lev)
{
    orc_level *F = &mg->lev[lev], *C = &mg->lev[lev + 1];
    for (i64 q = 0; q < C->n; q++) { C->msk[q] = 1; C->Rcoef[q] = 1.0; }
    for (i64 q = 0; q < F->n; q++) F->b[q] = (double)F->msk[q];
    restriction(mg, lev, 0);
    for (i64 q = 0; q < C->n; q++) C->msk[q] = (C->b[q] > 0.0) ? 1 : 0;
    apply_default_msk(mg, lev + 1);
}

/* core/mgfor/operators.f90:337-393 */
static void compute_Rcoef(orc_mg *mg, int lev)
{
    orc_level *F = &mg->lev[lev], *C = &mg->lev[lev + 1];
    for (i64 q = 0; q < C->n; q++) C->Rcoef[q] = (double)C->msk[q];
    for (i64 q = 0; q < F->n; q++) F->b[q] = 1.0;
    restriction(mg, lev, 0);
    for (i64 q = 0; q < C->n; q++) C->Rcoef[q] = 0.0;
    for (i64 q = 0; q < C->n; q++) if (C->b[q] > 0.0) C->y[q] = 4.0 / C->b[q];
    for (i64 q = 0; q < C->n; q++) C->y[q] = (double)C->msk[q] * C->y[q];
    for (i64 q = 0; q < C->n; q++) C->Rcoef[q] = C->y[q];
}

/* core/mgfor/operators.f90:395-424 */
static void compute_Pcoef(orc_mg *mg, int lev)
{
    orc_level *F = &mg->lev[lev], *C = &mg->lev[lev + 1];
    memset(F->x, 0, sizeof(double) * (size_t)F->n);
    for (i64 q = 0; q < C->n; q++) C->x[q] = (double)C->msk[q];
    for (i64 q = 0; q < F->n; q++) F->Pcoef[q] = (double)F->msk[q];
    prolongation(mg, lev);
    for (i64 q = 0; q < F->n; q++) F->Pcoef[q] = 0.0;
    for (i64 q = 0; q < F->n; q++) if (F->x[q] > 0.0) F->Pcoef[q] = 1.0 / F->x[q];
    for (i64 q = 0; q < F->n; q++) F->Pcoef[q] = (double)F->msk[q] * F->Pcoef[q];
}

/* core/mgfor/operators.f90:426-457 */
static void compute_diag(orc_mg *mg, int lev)
{
    orc_level *L = &mg->lev[lev];
    for (i64 q = 0; q < L->n; q++) L->x[q] = (double)L->msk[q];
    memset(L->b, 0, sizeof(double) * (size_t)L->n);
    memset(L->diag, 0, sizeof(double) * (size_t)L->n);
    residual(mg, lev);
    for (i64 q = 0; q < L->n; q++) L->diag[q] = -L->r[q];
    for (i64 q = 0; q < L->n; q++) L->idiag[q] = (L->diag[q] > 0.0) ? 1.0 / L->diag[q] : 0.0;
    memset(L->x, 0, sizeof(double) * (size_t)L->n);
}

/* core/mgfor/mg_setup.f90:225-307 (single rank: no glue).  Returns the number
 * of levels, or -1 if the coarsening recurrence never reaches x==2 or y==2. */
int orc_mg_hierarchy(int nx, int ny, int nz, int nh, int *lx, int *ly, int *lz, int maxlev)
{
    int x = nx, y = ny, z = nz + 2 * nh, i = 0;
    lx[0] = x; ly[0] = y; lz[0] = z;
    for (;;) {
        if (x == 2 || y == 2) break;
        x = x / 2; y = y / 2; z = z / 2 + nh;
        i++;
        if (i >= maxlev || x <= 0 || y <= 0) return -1;
        lx[i] = x; ly[i] = y; lz[i] = z;
    }
    return i + 1;
}

/* core/mgfor/mg_setup.f90:318-437.  nz is the interior count; the level-1
 * array z extent is nz+2nh (is3d). */
orc_mg *orc_mg_create(int nx, int ny, int nz, int topology, int maxite, double tol, double omega)
{
    const int nh = 3;                /* MG_Param%nh, core/mgfor/mg_types.f90:19 */
    int lx[50], ly[50], lz[50];
    int nl = orc_mg_hierarchy(nx, ny, nz, nh, lx, ly, lz, 50);
    if (nl < 1) return NULL;
    orc_mg *mg = (orc_mg *)calloc(1, sizeof(orc_mg));
    mg->nlevels = nl;
    mg->topology = topology;
    mg->xperio = (topology == T_XPERIO || topology == T_XYPERIO || topology == T_XYZPERIO);
    mg->yperio = (topology == T_YPERIO || topology == T_XYPERIO || topology == T_XYZPERIO);
    mg->zperio = (topology == T_ZPERIO || topology == T_XYZPERIO);
    mg->nh = nh;
    mg->maxite = maxite;
    mg->tol = tol;
    mg->omega = omega;
    i64 nmax = 0;
    for (int l = 0; l < nl; l++) {
        alloc_level(&mg->lev[l], lx[l], ly[l], lz[l], nh);
        if (mg->lev[l].n > nmax) nmax = mg->lev[l].n;
    }
    mg->tmp = (double *)malloc(sizeof(double) * (size_t)nmax);

    /* set_default_msk, core/mgfor/mg_setup.f90:180-211 */
    for (int l = 0; l < nl; l++) {
        orc_level *L = &mg->lev[l];
        for (int k = 1; k <= L->nz; k++)
            for (int j = 1; j <= L->ny; j++)
                for (int i = 1; i <= L->nx; i++)
                    if (mg->zperio || (k >= 1 + nh && k <= L->nz - nh))
                        L->msk[MI(L, i, j, k)] = 1;
    }
    /* setup_fine_msk, core/mgfor/mg_setup.f90:213-223 */
    fill_i(mg, &mg->lev[0], mg->lev[0].msk);
    /* setup_operators, core/mgfor/operators.f90:461-505 */
    for (int l = 0; l < nl - 1; l++) compute_msk(mg, l);
    for (int l = 0; l < nl - 1; l++) { compute_Rcoef(mg, l); compute_Pcoef(mg, l); }
    for (int l = 0; l < nl; l++) compute_diag(mg, l);
    return mg;
}

void orc_mg_destroy(orc_mg *mg)
{
    if (!mg) return;
    for (int l = 0; l < mg->nlevels; l++) {
        orc_level *L = &mg->lev[l];
        free(L->x); free(L->b); free(L->r); free(L->y);
        free(L->diag); free(L->idiag); free(L->Rcoef); free(L->Pcoef); free(L->msk);
    }
    free(mg->tmp);
    free(mg);
}

int orc_mg_nlevels(orc_mg *mg) { return mg->nlevels; }
void orc_mg_shape(orc_mg *mg, int lev, int *shape3)
{
    /* numpy shape (nz, ny+2nh, nx+2nh), as get_pyshape (core/mgfor/pytools.f90:8-15) */
    orc_level *L = &mg->lev[lev];
    shape3[0] = L->nz; shape3[1] = L->ny + 2 * L->nh; shape3[2] = L->nx + 2 * L->nh;
}
/* ivar numbering of core/mgfor/pytools.f90:17-42 */
double *orc_mg_array(orc_mg *mg, int lev, int ivar)
{
    orc_level *L = &mg->lev[lev];
    switch (ivar) {
    case 1: return L->x;
    case 2: return L->b;
    case 3: return L->r;
    case 4: return L->y;
    case 5: return L->diag;
    case 6: return L->idiag;
    case 8: return L->Rcoef;
    case 9: return L->Pcoef;
    default: return NULL;
    }
}
int *orc_mg_msk(orc_mg *mg, int lev) { return mg->lev[lev].msk; }
int orc_mg_last_nite(orc_mg *mg) { return mg->nite; }
double orc_mg_last_res(orc_mg *mg) { return mg->res; }
double orc_mg_last_normb(orc_mg *mg) { return mg->normb; }
void orc_mg_set_params(orc_mg *mg, int maxite, double tol, double omega)
{
    mg->maxite = maxite; mg->tol = tol; mg->omega = omega;
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
void orc_set_num_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}
