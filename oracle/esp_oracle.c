/*
 * esp_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU FP64 restatement of the ESP long-range electrostatics path, used as the
 * checker for the HIP implementation.  Nothing in the product links or calls
 * this file; only tests/, the smoke check and the CPU-baseline leg of the
 * benchmark load it.
 *
 * Contents
 *   oracle_split_eval   S(r)/S'(r) table lookup and L, -dL/dr
 *                       (EvaluatorPairPSWF.h:98-143, ESPForceCompute.cc:41-50,258-296)
 *   oracle_fft3d        mixed-radix complex 3-D FFT (x fastest), forward sign e^{-i}
 *   oracle_ewald        FP64 direct Ewald sum in 1/(4 pi r) units (Gaussian split)
 *   oracle_esp          the full ESP evaluation: spread -> FFT -> influence -> IFFT ->
 *                       analytic-derivative gather, pair correction, exclusions, self and
 *                       background terms, energy and virial.  Stage order follows
 *                       ESPForceCompute::computeForces (ESPForceCompute.cc:96-127); the stubs there
 *                       are filled from the mathematics of the split.
 *
 * Conventions kept from the reference: ESP_INV_4PI (ESPForceCompute.cc:36), mesh index
 * (cz*Ny + cy)*Nx + cx (ESPForceComputeGPU.cu:267), k ordering i <= N/2 -> +i else i - N
 * (ESPForceComputeGPU.cu:173-175), per-particle virial SoA xx,xy,xz,yy,yz,zz with a pitch
 * (ESPForceComputeGPU.cu:453-458,469-470), exclusion list ex_list[i*pitch + ex] with the
 * excluded-pair energy/force/virial formulas of ESPForceComputeGPU.cu:441-458.
 *
 * Parity status: every function here is pinned by tests (golden table vectors from
 * PSWF_Coeffs.h, numpy.fft, Madelung constants, two Ewald splitting parameters, the
 * Gaussian-split special case against Ewald, finite differences and brute force).
 */
#include <complex.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORACLE_INV_4PI 0.079577471545947673
#define ORACLE_PI 3.14159265358979323846

/* ------------------------------------------------------------------------- */
/* Table lookup: EvaluatorPairPSWF.h:113-143 with the reduced variable s = r/rc */
/* ------------------------------------------------------------------------- */
static double horner5(const double* coeffs, unsigned int seg, double t)
{
    /* ESPForceCompute.cc:41-50 */
    const double* c = coeffs + (size_t)seg * 5u;
    double v = c[4];
    v = c[3] + t * v;
    v = c[2] + t * v;
    v = c[1] + t * v;
    v = c[0] + t * v;
    return v;
}

static void lookup(const double* screen, const double* dscreen, const double* r_lo, const double* dr_inv,
                   unsigned int n_segs, double rc, double r, double* L, double* negdLdr, double* S, double* dS)
{
    if (r < 1.0e-12 || r > rc) {
        *L = 0.0;
        *negdLdr = 0.0;
        if (S) *S = 0.0;
        if (dS) *dS = 0.0;
        return;
    }
    /* segment from the clamped reduced distance, offset from the unclamped one
     * (EvaluatorPairPSWF.h:126-133 uses the unclamped r in t) */
    const double su = r / rc;
    double s = su;
    s = s < 0.0 ? 0.0 : (s > 1.0 - 1.0e-9 ? 1.0 - 1.0e-9 : s);
    unsigned int seg = (unsigned int)(s * (double)n_segs);
    if (seg >= n_segs) seg = n_segs - 1u;
    const double t = (su - r_lo[seg]) * dr_inv[seg];
    const double Sv = horner5(screen, seg, t);
    const double dSv = horner5(dscreen, seg, t) / rc; /* dS/dr from dS/ds */
    const double inv_4pi_r = ORACLE_INV_4PI / r;
    const double inv_4pi_r2 = inv_4pi_r / r;
    *L = Sv * inv_4pi_r;
    *negdLdr = Sv * inv_4pi_r2 - dSv * inv_4pi_r;
    if (S) *S = Sv;
    if (dS) *dS = dSv;
}

void oracle_split_eval(const double* screen, const double* dscreen, const double* r_lo, const double* dr_inv,
                       unsigned int n_segs, double rc, int n, const double* r, double* L, double* negdLdr,
                       double* S, double* dS)
{
    for (int i = 0; i < n; ++i)
        lookup(screen, dscreen, r_lo, dr_inv, n_segs, rc, r[i], &L[i], &negdLdr[i], S ? &S[i] : NULL,
               dS ? &dS[i] : NULL);
}

/* ------------------------------------------------------------------------- */
/* Mixed-radix FFT (recursive decimation in time over the smallest prime)     */
/* ------------------------------------------------------------------------- */
static int smallest_factor(int n)
{
    for (int p = 2; p * p <= n; ++p)
        if (n % p == 0) return p;
    return n;
}

static void fft_rec(int n, int stride, const double complex* in, double complex* out, int sign, double complex* tmp)
{
    if (n == 1) {
        out[0] = in[0];
        return;
    }
    const int p = smallest_factor(n);
    const int m = n / p;
    if (p == n) { /* prime: direct DFT */
        for (int k = 0; k < n; ++k) {
            double complex s = 0;
            for (int j = 0; j < n; ++j) {
                const double ang = sign * 2.0 * ORACLE_PI * (double)((long)j * k % n) / (double)n;
                s += in[(size_t)j * stride] * (cos(ang) + I * sin(ang));
            }
            out[k] = s;
        }
        return;
    }
    /* p sub-transforms of length m on in[r*stride + j*p*stride] */
    for (int r = 0; r < p; ++r)
        fft_rec(m, stride * p, in + (size_t)r * stride, tmp + (size_t)r * m, sign, out);
    for (int k = 0; k < m; ++k) {
        for (int q = 0; q < p; ++q) {
            const int kk = k + q * m;
            double complex s = 0;
            for (int r = 0; r < p; ++r) {
                const double ang = sign * 2.0 * ORACLE_PI * (double)((long)r * kk % n) / (double)n;
                s += tmp[(size_t)r * m + k] * (cos(ang) + I * sin(ang));
            }
            out[kk] = s;
        }
    }
}

static void fft1d(int n, double complex* line, int sign)
{
    double complex* in = malloc(sizeof(double complex) * n);
    double complex* tmp = malloc(sizeof(double complex) * n);
    memcpy(in, line, sizeof(double complex) * n);
    fft_rec(n, 1, in, line, sign, tmp);
    free(in);
    free(tmp);
}

/* data: complex interleaved, index (z*ny + y)*nx + x; sign -1 forward, +1 backward (unnormalised) */
int oracle_fft3d(int nx, int ny, int nz, double* data, int sign)
{
    double complex* a = (double complex*)data;
    int nmax = nx > ny ? nx : ny;
    nmax = nmax > nz ? nmax : nz;
    double complex* line = malloc(sizeof(double complex) * nmax);
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y) {
            double complex* p = a + ((size_t)z * ny + y) * nx;
            fft1d(nx, p, sign);
        }
    for (int z = 0; z < nz; ++z)
        for (int x = 0; x < nx; ++x) {
            for (int y = 0; y < ny; ++y) line[y] = a[((size_t)z * ny + y) * nx + x];
            fft1d(ny, line, sign);
            for (int y = 0; y < ny; ++y) a[((size_t)z * ny + y) * nx + x] = line[y];
        }
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
            for (int z = 0; z < nz; ++z) line[z] = a[((size_t)z * ny + y) * nx + x];
            fft1d(nz, line, sign);
            for (int z = 0; z < nz; ++z) a[((size_t)z * ny + y) * nx + x] = line[z];
        }
    free(line);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Cell binning for real-space sums (any cutoff, periodic images included)    */
/* ------------------------------------------------------------------------- */
/* General (triclinic) box of HOOMD's BoxDim, centred on the origin: lattice vectors
 * a1 = (Lx, 0, 0), a2 = (xy Ly, Ly, 0), a3 = (xz Lz, yz Lz, Lz). */
typedef struct {
    double L[3];
    double xy, xz, yz;
} obox_t;

/* fractions u = B^-1 r + 1/2 (BoxDim::makeFraction), not wrapped */
static void box_fraction3(const obox_t* b, const double* r, double* u)
{
    u[0] = (r[0] - b->xy * r[1] - (b->xz - b->xy * b->yz) * r[2]) / b->L[0] + 0.5;
    u[1] = (r[1] - b->yz * r[2]) / b->L[1] + 0.5;
    u[2] = r[2] / b->L[2] + 0.5;
}

/* BoxDim::minImage: a3 by rint(d_z / Lz), then a2 by rint(d_y / Ly), then a1 by rint(d_x / Lx);
 * the nearest image whenever |d| is below half the smallest perpendicular width of the box */
static void min_image_box(double* d, const obox_t* b)
{
    const double iz = nearbyint(d[2] / b->L[2]);
    d[2] -= b->L[2] * iz;
    d[1] -= b->yz * b->L[2] * iz;
    d[0] -= b->xz * b->L[2] * iz;
    const double iy = nearbyint(d[1] / b->L[1]);
    d[1] -= b->L[1] * iy;
    d[0] -= b->xy * b->L[1] * iy;
    d[0] -= b->L[0] * nearbyint(d[0] / b->L[0]);
}

/* k = 2 pi B^-T m for integer mode m */
static void box_wavevector(const obox_t* b, int mx, int my, int mz, double* k)
{
    const double k0x = 2 * ORACLE_PI * mx / b->L[0], k0y = 2 * ORACLE_PI * my / b->L[1], k0z = 2 * ORACLE_PI * mz / b->L[2];
    k[0] = k0x;
    k[1] = k0y - b->xy * k0x;
    k[2] = k0z - b->yz * k0y + (b->xy * b->yz - b->xz) * k0x;
}

/* half the smallest perpendicular width of the box: 1 / (2 max |b_a|), b_a the rows of B^-1 */
static double box_half_width(const obox_t* b)
{
    const double n1 = sqrt(1.0 + b->xy * b->xy + (b->xy * b->yz - b->xz) * (b->xy * b->yz - b->xz)) / b->L[0];
    const double n2 = sqrt(1.0 + b->yz * b->yz) / b->L[1];
    const double n3 = 1.0 / b->L[2];
    double m = n1 > n2 ? n1 : n2;
    m = m > n3 ? m : n3;
    return 0.5 / m;
}

typedef struct {
    int nc[3];       /* cells per axis */
    int reach[3];    /* cell offsets to visit per axis: |offset| <= reach */
    int* start;      /* [ncells + 1] */
    int* idx;        /* particle indices ordered by cell (increasing index within a cell) */
    double* xyz;     /* their coordinates wrapped into the primary box, in the same order */
} cells_t;

/* cell of coordinate x (box centred on the origin) along an axis of length L with nc cells */
static int cell_of(double x, double L, int nc)
{
    double f = x / L + 0.5;
    f -= floor(f);
    int c = (int)(f * nc);
    return c >= nc ? nc - 1 : (c < 0 ? 0 : c);
}

/* Bin N particles into cells of edge >= rcut / 3 (at most 256 per axis).  A source image within
 * rcut of a target lies at most `reach` cells away along each axis, counting cells across the
 * periodic boundary as images (cell c + s*nc is cell c shifted by s box lengths). */
static void cells_build(cells_t* c, int N, const double* postype, const double* L, double rcut)
{
    size_t ncell = 1;
    for (int a = 0; a < 3; ++a) {
        int n = (int)floor(L[a] / (rcut / 3.0));
        if (n < 1) n = 1;
        if (n > 256) n = 256;
        c->nc[a] = n;
        c->reach[a] = (int)ceil(rcut / (L[a] / n));
        ncell *= (size_t)n;
    }
    c->start = calloc(ncell + 1, sizeof(int));
    c->idx = malloc(sizeof(int) * (size_t)(N > 0 ? N : 1));
    int* cell = malloc(sizeof(int) * (size_t)(N > 0 ? N : 1));
    for (int j = 0; j < N; ++j) {
        const int cx = cell_of(postype[4 * j], L[0], c->nc[0]);
        const int cy = cell_of(postype[4 * j + 1], L[1], c->nc[1]);
        const int cz = cell_of(postype[4 * j + 2], L[2], c->nc[2]);
        cell[j] = (cz * c->nc[1] + cy) * c->nc[0] + cx;
        c->start[cell[j] + 1]++;
    }
    for (size_t k = 0; k < ncell; ++k) c->start[k + 1] += c->start[k];
    int* fill = malloc(sizeof(int) * (ncell > 0 ? ncell : 1));
    memcpy(fill, c->start, sizeof(int) * ncell);
    for (int j = 0; j < N; ++j) c->idx[fill[cell[j]]++] = j;
    c->xyz = malloc(sizeof(double) * 3 * (size_t)(N > 0 ? N : 1));
    for (int k = 0; k < N; ++k)
        for (int a = 0; a < 3; ++a) {
            const double f = postype[4 * c->idx[k] + a] / L[a] + 0.5;
            c->xyz[3 * k + a] = (f - floor(f) - 0.5) * L[a];
        }
    free(fill);
    free(cell);
}

static void cells_free(cells_t* c)
{
    free(c->start);
    free(c->idx);
    free(c->xyz);
}

static int floor_div(int a, int n) { return (a >= 0) ? a / n : -((-a + n - 1) / n); }

/* Call fn(ctx, j, d) for every source image j (excluding i itself at zero shift) with
 * |d| < rcut, d = r_i - (r_j + shift).  Each (j, image) pair is visited once. */
typedef void (*pair_fn)(void* ctx, int j, const double* d, double r2);
static void cells_for_neighbours(const cells_t* c, const double* postype, const double* L, int i, const double* ri,
                                 double rcut, pair_fn fn, void* ctx)
{
    const double rc2 = rcut * rcut;
    const int ci[3] = {cell_of(ri[0], L[0], c->nc[0]), cell_of(ri[1], L[1], c->nc[1]), cell_of(ri[2], L[2], c->nc[2])};
    for (int oz = -c->reach[2]; oz <= c->reach[2]; ++oz) {
        const int vz = ci[2] + oz, wz = vz - floor_div(vz, c->nc[2]) * c->nc[2];
        const double sz = floor_div(vz, c->nc[2]) * L[2];
        for (int oy = -c->reach[1]; oy <= c->reach[1]; ++oy) {
            const int vy = ci[1] + oy, wy = vy - floor_div(vy, c->nc[1]) * c->nc[1];
            const double sy = floor_div(vy, c->nc[1]) * L[1];
            for (int ox = -c->reach[0]; ox <= c->reach[0]; ++ox) {
                const int vx = ci[0] + ox, wx = vx - floor_div(vx, c->nc[0]) * c->nc[0];
                const double sx = floor_div(vx, c->nc[0]) * L[0];
                const int zero_shift = (sx == 0.0 && sy == 0.0 && sz == 0.0);
                const int cell = (wz * c->nc[1] + wy) * c->nc[0] + wx;
                /* the source's coordinate is taken inside the primary box, as its cell was */
                const double px = ri[0] - sx, py = ri[1] - sy, pz = ri[2] - sz;
                for (int k = c->start[cell]; k < c->start[cell + 1]; ++k) {
                    const double d[3] = {px - c->xyz[3 * k], py - c->xyz[3 * k + 1], pz - c->xyz[3 * k + 2]};
                    const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
                    if (r2 >= rc2) continue;
                    const int j = c->idx[k];
                    if (j == i && zero_shift) continue;
                    fn(ctx, j, d, r2);
                }
            }
        }
    }
}

/* ------------------------------------------------------------------------- */
/* Direct Ewald sum (Gaussian split), orthorhombic box centred on the origin */
/* ------------------------------------------------------------------------- */
typedef struct {
    const double* q;
    double qi, alpha, two_a_spi;
    int want_ev;
    double f[3], E, W[6];
} ewald_real_ctx;

static void ewald_real_pair(void* vctx, int j, const double* d, double r2)
{
    ewald_real_ctx* c = (ewald_real_ctx*)vctx;
    const double r = sqrt(r2);
    const double qq = c->qi * c->q[j] * ORACLE_INV_4PI;
    const double er = erfc(c->alpha * r);
    const double fdivr = qq * (er / r + c->two_a_spi * exp(-c->alpha * c->alpha * r2)) / r2;
    for (int k = 0; k < 3; ++k) c->f[k] += fdivr * d[k];
    if (c->want_ev) {
        c->E += 0.5 * qq * er / r;
        c->W[0] += 0.5 * d[0] * d[0] * fdivr;
        c->W[1] += 0.5 * d[0] * d[1] * fdivr;
        c->W[2] += 0.5 * d[0] * d[2] * fdivr;
        c->W[3] += 0.5 * d[1] * d[1] * fdivr;
        c->W[4] += 0.5 * d[1] * d[2] * fdivr;
        c->W[5] += 0.5 * d[2] * d[2] * fdivr;
    }
}

/* one row of the reciprocal half space: fixed (b, c), a in [a_lo, a_hi], S stored from off */
typedef struct { int b, c, a_lo, a_hi; long off; } krow_t;

/*
 * Forces on the n_targets particles listed in targets (all if targets == NULL) from all N
 * particles; energy and virial of the whole system when energy != NULL (needs all targets).
 * Real space: every periodic image within rcut (rcut may exceed L/2), found through a cell
 * list.  k space: |k| <= kcut over the half space, the structure factor accumulated per
 * thread from per-particle phase factors e^{-i k_a x} (one sincos per axis wavenumber).
 */
int oracle_ewald(int N, const double* postype, const double* q, const double* Lbox, double alpha, double rcut,
                 double kcut, int n_targets, const int* targets, double* force, double* energy, double* virial)
{
    const double L[3] = {Lbox[0], Lbox[1], Lbox[2]};
    const double V = L[0] * L[1] * L[2];
    const int nt = targets ? n_targets : N;
    double E = 0.0, W[6] = {0, 0, 0, 0, 0, 0};
    const int want_ev = (energy != NULL) || (virial != NULL);
    const double two_a_spi = 2.0 * alpha / sqrt(ORACLE_PI);

    memset(force, 0, sizeof(double) * 3 * (size_t)nt);
    /* real space */
    cells_t cl;
    cells_build(&cl, N, postype, L, rcut);
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : E) reduction(+ : W[:6])
    for (int t = 0; t < nt; ++t) {
        const int i = targets ? targets[t] : t;
        ewald_real_ctx c;
        memset(&c, 0, sizeof c);
        c.q = q;
        c.qi = q[i];
        c.alpha = alpha;
        c.two_a_spi = two_a_spi;
        c.want_ev = want_ev;
        double ri[3];
        for (int a = 0; a < 3; ++a) {
            double f = postype[4 * i + a] / L[a] + 0.5;
            ri[a] = (f - floor(f) - 0.5) * L[a];
        }
        cells_for_neighbours(&cl, postype, L, i, ri, rcut, ewald_real_pair, &c);
        for (int k = 0; k < 3; ++k) force[3 * t + k] = c.f[k];
        E += c.E;
        for (int k = 0; k < 6; ++k) W[k] += c.W[k];
    }
    cells_free(&cl);

    /* reciprocal space over the half space k_z > 0 or (k_z == 0 and (k_y > 0 or (k_y == 0 and k_x > 0))) */
    const int mx = (int)floor(kcut * L[0] / (2 * ORACLE_PI)), my = (int)floor(kcut * L[1] / (2 * ORACLE_PI)),
              mz = (int)floor(kcut * L[2] / (2 * ORACLE_PI));
    const double kc2 = kcut * kcut;
    const double ux = 2 * ORACLE_PI / L[0], uy = 2 * ORACLE_PI / L[1], uz = 2 * ORACLE_PI / L[2];
    krow_t* rows = malloc(sizeof(krow_t) * (size_t)(mz + 1) * (2 * my + 1));
    long nrow = 0, nk = 0;
    for (int c = 0; c <= mz; ++c)
        for (int b = -my; b <= my; ++b) {
            if (c == 0 && b < 0) continue;
            const double rest = kc2 - (b * uy) * (b * uy) - (c * uz) * (c * uz);
            if (rest < 0) continue;
            int am = (int)floor(sqrt(rest) / ux);
            if (am > mx) am = mx;
            while (am >= 0 && (am * ux) * (am * ux) > rest) --am;
            if (am < 0) continue;
            int a_lo = -am;
            if (c == 0 && b == 0) a_lo = 1;
            if (a_lo > am) continue;
            rows[nrow].b = b;
            rows[nrow].c = c;
            rows[nrow].a_lo = a_lo;
            rows[nrow].a_hi = am;
            rows[nrow].off = nk;
            nk += am - a_lo + 1;
            ++nrow;
        }
    const size_t nkz = (size_t)(nk > 0 ? nk : 1);
    double* Sre = calloc(nkz, sizeof(double));
    double* Sim = calloc(nkz, sizeof(double));
    double* gk = malloc(sizeof(double) * nkz);
    for (long r = 0; r < nrow; ++r)
        for (int a = rows[r].a_lo; a <= rows[r].a_hi; ++a) {
            const double kx = a * ux, ky = rows[r].b * uy, kz = rows[r].c * uz;
            const double k2 = kx * kx + ky * ky + kz * kz;
            gk[rows[r].off + a - rows[r].a_lo] = exp(-k2 / (4 * alpha * alpha)) / k2;
        }
    const int wx = 2 * mx + 1, wy = 2 * my + 1;
#pragma omp parallel
    {
        double* sr = calloc(nkz, sizeof(double));
        double* si = calloc(nkz, sizeof(double));
        double* ex = malloc(sizeof(double) * 2 * wx);
        double* ey = malloc(sizeof(double) * 2 * wy);
        double* ez = malloc(sizeof(double) * 2 * (mz + 1));
#pragma omp for schedule(static)
        for (int j = 0; j < N; ++j) {
            /* e^{-i k x_j} per axis wavenumber */
            for (int a = -mx; a <= mx; ++a) {
                ex[2 * (a + mx)] = cos(a * ux * postype[4 * j]);
                ex[2 * (a + mx) + 1] = -sin(a * ux * postype[4 * j]);
            }
            for (int b = -my; b <= my; ++b) {
                ey[2 * (b + my)] = cos(b * uy * postype[4 * j + 1]);
                ey[2 * (b + my) + 1] = -sin(b * uy * postype[4 * j + 1]);
            }
            for (int c = 0; c <= mz; ++c) {
                ez[2 * c] = cos(c * uz * postype[4 * j + 2]);
                ez[2 * c + 1] = -sin(c * uz * postype[4 * j + 2]);
            }
            for (long r = 0; r < nrow; ++r) {
                const double* yb = &ey[2 * (rows[r].b + my)];
                const double* zc = &ez[2 * rows[r].c];
                const double tr = q[j] * (yb[0] * zc[0] - yb[1] * zc[1]);
                const double ti = q[j] * (yb[0] * zc[1] + yb[1] * zc[0]);
                double* Sr = sr + rows[r].off - rows[r].a_lo;
                double* Si = si + rows[r].off - rows[r].a_lo;
                for (int a = rows[r].a_lo; a <= rows[r].a_hi; ++a) {
                    const double xr = ex[2 * (a + mx)], xi = ex[2 * (a + mx) + 1];
                    Sr[a] += tr * xr - ti * xi;
                    Si[a] += tr * xi + ti * xr;
                }
            }
        }
#pragma omp critical
        for (size_t k = 0; k < (size_t)nk; ++k) {
            Sre[k] += sr[k];
            Sim[k] += si[k];
        }
        free(sr);
        free(si);
        free(ex);
        free(ey);
        free(ez);
    }
    double Ek = 0.0, Wk[6] = {0, 0, 0, 0, 0, 0};
    if (want_ev) {
        for (long r = 0; r < nrow; ++r)
            for (int a = rows[r].a_lo; a <= rows[r].a_hi; ++a) {
                const long k = rows[r].off + a - rows[r].a_lo;
                const double kx = a * ux, ky = rows[r].b * uy, kz = rows[r].c * uz;
                const double k2 = kx * kx + ky * ky + kz * kz;
                const double s2 = Sre[k] * Sre[k] + Sim[k] * Sim[k];
                /* half space: factor 2 */
                const double e = gk[k] * s2 / V;
                Ek += e;
                const double fac = -2.0 / k2 * (1.0 + k2 / (4 * alpha * alpha));
                Wk[0] += e * (1.0 + fac * kx * kx);
                Wk[1] += e * (fac * kx * ky);
                Wk[2] += e * (fac * kx * kz);
                Wk[3] += e * (1.0 + fac * ky * ky);
                Wk[4] += e * (fac * ky * kz);
                Wk[5] += e * (1.0 + fac * kz * kz);
            }
    }
#pragma omp parallel
    {
        double* ex = malloc(sizeof(double) * 2 * wx);
        double* ey = malloc(sizeof(double) * 2 * wy);
        double* ez = malloc(sizeof(double) * 2 * (mz + 1));
#pragma omp for schedule(dynamic, 16)
        for (int t = 0; t < nt; ++t) {
            const int j = targets ? targets[t] : t;
            /* e^{+i k x_j} per axis wavenumber */
            for (int a = -mx; a <= mx; ++a) {
                ex[2 * (a + mx)] = cos(a * ux * postype[4 * j]);
                ex[2 * (a + mx) + 1] = sin(a * ux * postype[4 * j]);
            }
            for (int b = -my; b <= my; ++b) {
                ey[2 * (b + my)] = cos(b * uy * postype[4 * j + 1]);
                ey[2 * (b + my) + 1] = sin(b * uy * postype[4 * j + 1]);
            }
            for (int c = 0; c <= mz; ++c) {
                ez[2 * c] = cos(c * uz * postype[4 * j + 2]);
                ez[2 * c + 1] = sin(c * uz * postype[4 * j + 2]);
            }
            double f[3] = {0, 0, 0};
            for (long r = 0; r < nrow; ++r) {
                const double* yb = &ey[2 * (rows[r].b + my)];
                const double* zc = &ez[2 * rows[r].c];
                const double tr = yb[0] * zc[0] - yb[1] * zc[1];
                const double ti = yb[0] * zc[1] + yb[1] * zc[0];
                double gx = 0, gyz = 0;
                for (int a = rows[r].a_lo; a <= rows[r].a_hi; ++a) {
                    const long k = rows[r].off + a - rows[r].a_lo;
                    const double xr = ex[2 * (a + mx)], xi = ex[2 * (a + mx) + 1];
                    /* Im(e^{i k r_j} S(k)) weighted by G(k) */
                    const double er = tr * xr - ti * xi, ei = tr * xi + ti * xr;
                    const double im = gk[k] * (ei * Sre[k] + er * Sim[k]);
                    gx += im * (a * ux);
                    gyz += im;
                }
                f[0] += gx;
                f[1] += gyz * (rows[r].b * uy);
                f[2] += gyz * (rows[r].c * uz);
            }
            /* times 2 for the half space */
            const double s = 2.0 * q[j] / V;
            for (int k = 0; k < 3; ++k) force[3 * t + k] += s * f[k];
        }
        free(ex);
        free(ey);
        free(ez);
    }
    if (want_ev) {
        double q2 = 0, qs = 0;
        for (int j = 0; j < N; ++j) {
            q2 += q[j] * q[j];
            qs += q[j];
        }
        const double Eself = -alpha / sqrt(ORACLE_PI) * ORACLE_INV_4PI * q2;
        const double Ebg = -qs * qs / (8.0 * V * alpha * alpha);
        if (energy) *energy = E + Ek + Eself + Ebg;
        if (virial) {
            for (int k = 0; k < 6; ++k) virial[k] = W[k] + Wk[k];
            /* background term: E_bg scales as 1/V -> -dE/d eps_aa = E_bg */
            virial[0] += Ebg;
            virial[3] += Ebg;
            virial[5] += Ebg;
        }
    }
    free(rows);
    free(gk);
    free(Sre);
    free(Sim);
    return 0;
}

/*
 * Direct Ewald sum in a general (triclinic) box: real space over the nearest image of every
 * pair (rcut must be below half the smallest perpendicular width of the box), k space over
 * k = 2 pi B^-T m with |k| <= kcut on the half space of m, the structure factor from per-axis
 * phase factors e^{-2 pi i m_a u_a} of the fractions u = B^-1 r.  Forces on all particles,
 * energy and virial.  An independent check of the mesh method in tilted boxes.
 */
int oracle_ewald_box(int N, const double* postype, const double* q, const double* box6, double alpha, double rcut, double kcut,
                     double* force, double* energy, double* virial)
{
    obox_t b = {{box6[0], box6[1], box6[2]}, box6[3], box6[4], box6[5]};
    if (rcut >= box_half_width(&b)) return -1;
    const double V = b.L[0] * b.L[1] * b.L[2];
    const double two_a_spi = 2.0 * alpha / sqrt(ORACLE_PI);
    const double rc2 = rcut * rcut;
    double E = 0.0, W[6] = {0, 0, 0, 0, 0, 0};
    memset(force, 0, sizeof(double) * 3 * (size_t)N);
#pragma omp parallel for schedule(dynamic, 8) reduction(+ : E) reduction(+ : W[:6])
    for (int i = 0; i < N; ++i) {
        double f[3] = {0, 0, 0};
        for (int j = 0; j < N; ++j) {
            if (j == i) continue;
            double d[3] = {postype[4 * i] - postype[4 * j], postype[4 * i + 1] - postype[4 * j + 1],
                           postype[4 * i + 2] - postype[4 * j + 2]};
            min_image_box(d, &b);
            const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            if (r2 >= rc2) continue;
            const double r = sqrt(r2);
            const double qq = q[i] * q[j] * ORACLE_INV_4PI;
            const double er = erfc(alpha * r);
            const double fdivr = qq * (er / r + two_a_spi * exp(-alpha * alpha * r2)) / r2;
            for (int a = 0; a < 3; ++a) f[a] += fdivr * d[a];
            E += 0.5 * qq * er / r;
            W[0] += 0.5 * d[0] * d[0] * fdivr;
            W[1] += 0.5 * d[0] * d[1] * fdivr;
            W[2] += 0.5 * d[0] * d[2] * fdivr;
            W[3] += 0.5 * d[1] * d[1] * fdivr;
            W[4] += 0.5 * d[1] * d[2] * fdivr;
            W[5] += 0.5 * d[2] * d[2] * fdivr;
        }
        for (int a = 0; a < 3; ++a) force[3 * i + a] = f[a];
    }
    /* |m_a| <= |a_a| kcut / (2 pi): m_a = a_a . k / (2 pi) */
    const double alen[3] = {b.L[0], b.L[1] * sqrt(1.0 + b.xy * b.xy), b.L[2] * sqrt(1.0 + b.xz * b.xz + b.yz * b.yz)};
    int mm[3];
    for (int a = 0; a < 3; ++a) mm[a] = (int)floor(alen[a] * kcut / (2 * ORACLE_PI));
    const int wx = 2 * mm[0] + 1, wy = 2 * mm[1] + 1, wz = mm[2] + 1;
    /* mode list on the half space m_z > 0, or m_z = 0 and (m_y > 0 or (m_y = 0 and m_x > 0)) */
    long nk = 0;
    int* modes = malloc(sizeof(int) * 3 * (size_t)wx * wy * wz);
    double* kv = malloc(sizeof(double) * 3 * (size_t)wx * wy * wz);
    for (int c = 0; c <= mm[2]; ++c)
        for (int bb = -mm[1]; bb <= mm[1]; ++bb)
            for (int a = -mm[0]; a <= mm[0]; ++a) {
                if (c == 0 && (bb < 0 || (bb == 0 && a <= 0))) continue;
                double k[3];
                box_wavevector(&b, a, bb, c, k);
                if (k[0] * k[0] + k[1] * k[1] + k[2] * k[2] > kcut * kcut) continue;
                modes[3 * nk] = a;
                modes[3 * nk + 1] = bb;
                modes[3 * nk + 2] = c;
                for (int t = 0; t < 3; ++t) kv[3 * nk + t] = k[t];
                ++nk;
            }
    const size_t nkz = (size_t)(nk > 0 ? nk : 1);
    double* Sre = calloc(nkz, sizeof(double));
    double* Sim = calloc(nkz, sizeof(double));
    double* ph = malloc(sizeof(double) * 2 * (size_t)(wx + wy + wz) * (size_t)(N > 0 ? N : 1));
    /* e^{-2 pi i m_a u_a} per particle and axis wavenumber */
    for (int j = 0; j < N; ++j) {
        double u[3];
        box_fraction3(&b, &postype[4 * j], u);
        double* pj = ph + 2 * (size_t)(wx + wy + wz) * j;
        for (int a = -mm[0]; a <= mm[0]; ++a) {
            pj[2 * (a + mm[0])] = cos(2 * ORACLE_PI * a * u[0]);
            pj[2 * (a + mm[0]) + 1] = -sin(2 * ORACLE_PI * a * u[0]);
        }
        for (int bb = -mm[1]; bb <= mm[1]; ++bb) {
            pj[2 * (wx + bb + mm[1])] = cos(2 * ORACLE_PI * bb * u[1]);
            pj[2 * (wx + bb + mm[1]) + 1] = -sin(2 * ORACLE_PI * bb * u[1]);
        }
        for (int c = 0; c <= mm[2]; ++c) {
            pj[2 * (wx + wy + c)] = cos(2 * ORACLE_PI * c * u[2]);
            pj[2 * (wx + wy + c) + 1] = -sin(2 * ORACLE_PI * c * u[2]);
        }
    }
#pragma omp parallel for schedule(static)
    for (long k = 0; k < nk; ++k) {
        double sr = 0, si = 0;
        for (int j = 0; j < N; ++j) {
            const double* pj = ph + 2 * (size_t)(wx + wy + wz) * j;
            const double* ex = &pj[2 * (modes[3 * k] + mm[0])];
            const double* ey = &pj[2 * (wx + modes[3 * k + 1] + mm[1])];
            const double* ez = &pj[2 * (wx + wy + modes[3 * k + 2])];
            const double tr = ey[0] * ez[0] - ey[1] * ez[1], ti = ey[0] * ez[1] + ey[1] * ez[0];
            sr += q[j] * (tr * ex[0] - ti * ex[1]);
            si += q[j] * (tr * ex[1] + ti * ex[0]);
        }
        Sre[k] = sr;
        Sim[k] = si;
    }
    double Ek = 0.0, Wk[6] = {0, 0, 0, 0, 0, 0};
    double* gk = malloc(sizeof(double) * nkz);
    for (long k = 0; k < nk; ++k) {
        const double* kk = &kv[3 * k];
        const double k2 = kk[0] * kk[0] + kk[1] * kk[1] + kk[2] * kk[2];
        gk[k] = exp(-k2 / (4 * alpha * alpha)) / k2;
        /* half space: factor 2 of 1 / (2V) */
        const double e = gk[k] * (Sre[k] * Sre[k] + Sim[k] * Sim[k]) / V;
        Ek += e;
        const double fac = -2.0 / k2 * (1.0 + k2 / (4 * alpha * alpha));
        Wk[0] += e * (1.0 + fac * kk[0] * kk[0]);
        Wk[1] += e * (fac * kk[0] * kk[1]);
        Wk[2] += e * (fac * kk[0] * kk[2]);
        Wk[3] += e * (1.0 + fac * kk[1] * kk[1]);
        Wk[4] += e * (fac * kk[1] * kk[2]);
        Wk[5] += e * (1.0 + fac * kk[2] * kk[2]);
    }
#pragma omp parallel for schedule(static)
    for (int j = 0; j < N; ++j) {
        const double* pj = ph + 2 * (size_t)(wx + wy + wz) * j;
        double f[3] = {0, 0, 0};
        for (long k = 0; k < nk; ++k) {
            const double* ex = &pj[2 * (modes[3 * k] + mm[0])];
            const double* ey = &pj[2 * (wx + modes[3 * k + 1] + mm[1])];
            const double* ez = &pj[2 * (wx + wy + modes[3 * k + 2])];
            /* e^{+i k r_j} = conj of the stored phases */
            const double tr = ey[0] * ez[0] - ey[1] * ez[1], ti = ey[0] * ez[1] + ey[1] * ez[0];
            const double er = tr * ex[0] - ti * ex[1], ei = -(tr * ex[1] + ti * ex[0]);
            const double im = gk[k] * (ei * Sre[k] + er * Sim[k]);
            for (int a = 0; a < 3; ++a) f[a] += im * kv[3 * k + a];
        }
        const double sc = 2.0 * q[j] / V;
        for (int a = 0; a < 3; ++a) force[3 * j + a] += sc * f[a];
    }
    double q2 = 0, qs = 0;
    for (int j = 0; j < N; ++j) {
        q2 += q[j] * q[j];
        qs += q[j];
    }
    const double Eself = -alpha / sqrt(ORACLE_PI) * ORACLE_INV_4PI * q2;
    const double Ebg = -qs * qs / (8.0 * V * alpha * alpha);
    if (energy) *energy = E + Ek + Eself + Ebg;
    if (virial) {
        for (int k = 0; k < 6; ++k) virial[k] = W[k] + Wk[k];
        virial[0] += Ebg;
        virial[3] += Ebg;
        virial[5] += Ebg;
    }
    free(modes);
    free(kv);
    free(Sre);
    free(Sim);
    free(ph);
    free(gk);
    return 0;
}

/* OpenMP threads the oracle's parallel loops use */
int oracle_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------- */
/* Full ESP evaluation                                                        */
/* ------------------------------------------------------------------------- */
typedef struct {
    int N;
    const double* postype; /* Scalar4[N], box centred on the origin */
    const double* q;
    double L[3];
    int n[3];              /* mesh Nx, Ny, Nz */
    int P;                 /* stencil order */
    int D;                 /* window degree */
    const double* wcoef;   /* [P][D+1] monomials in z = 2t - 1 */
    int n_segs;
    const double* screen;  /* [n_segs][5] S(s) */
    const double* dscreen; /* [n_segs][5] dS/ds */
    const double* r_lo;
    const double* dr_inv;
    double rc;
    int split_kind;        /* 0 PSWF (F from the table), 1 Gaussian erfc(kappa r) */
    double kappa;
    const unsigned int* n_ex; /* may be NULL */
    const unsigned int* ex_list;
    unsigned int ex_pitch;
    double k_cut;          /* spherical reciprocal-space cutoff: G = 0 for |k| > k_cut (0 = none) */
    double alpha;          /* screening: G(k) carries exp(-alpha |k|) (ESPForceComputeGPU.cu:199); 0 = none */
    double tilt[3];        /* HOOMD tilt factors xy, xz, yz (0 = orthorhombic) */
} oracle_esp_in;

typedef struct {
    double* force;        /* [N][4]  (fx, fy, fz, per-particle pair + exclusion energy) */
    double* virial;       /* [6][pitch] per-particle pair + exclusion virial */
    unsigned int virial_pitch;
    double ext_energy;    /* reciprocal + self + background */
    double ext_virial[6]; /* reciprocal + background */
    double* mesh;         /* optional [Nz][Ny][Nx] spread charge */
    double* phi;          /* optional [Nz][Ny][Nx] potential mesh */
    double* green;        /* optional [Nz][Ny][Nx] influence function G(k) (full spectrum) */
} oracle_esp_out;

static double window_eval(const double* c, int D, double t, double* deriv)
{
    const double z = 2.0 * t - 1.0;
    double v = c[D], dv = 0.0;
    for (int j = D - 1; j >= 0; --j) {
        dv = dv * z + v;
        v = v * z + c[j];
    }
    if (deriv) *deriv = 2.0 * dv; /* d/dt = 2 d/dz */
    return v;
}

/* |int W(d) e^{-i theta d} dd|^2 by 32-point Gauss per tap */
static double window_fourier2(const double* wc, int P, int D, double theta)
{
    static const int NG = 32;
    static double gx[32], gw[32];
    static int init = 0;
    if (!init) {
        /* Gauss-Legendre on [0,1] by Newton */
        for (int i = 0; i < NG; ++i) {
            double z = cos(ORACLE_PI * (i + 0.75) / (NG + 0.5));
            for (int it = 0; it < 100; ++it) {
                double p0 = 1, p1 = z;
                for (int k = 2; k <= NG; ++k) {
                    double p2 = ((2.0 * k - 1) * z * p1 - (k - 1.0) * p0) / k;
                    p0 = p1;
                    p1 = p2;
                }
                double dp = NG * (z * p1 - p0) / (z * z - 1);
                double dz = p1 / dp;
                z -= dz;
                if (fabs(dz) < 1e-16) {
                    p0 = 1;
                    p1 = z;
                    for (int k = 2; k <= NG; ++k) {
                        double p2 = ((2.0 * k - 1) * z * p1 - (k - 1.0) * p0) / k;
                        p0 = p1;
                        p1 = p2;
                    }
                    dp = NG * (z * p1 - p0) / (z * z - 1);
                    gx[i] = 0.5 * (1 - z);
                    gw[i] = 1.0 / ((1 - z * z) * dp * dp);
                    break;
                }
            }
        }
        init = 1;
    }
    double re = 0, im = 0;
    for (int k = 0; k < P; ++k)
        for (int g = 0; g < NG; ++g) {
            const double t = gx[g];
            const double d = k - 0.5 * P + t;
            const double w = gw[g] * window_eval(wc + (size_t)k * (D + 1), D, t, NULL);
            re += w * cos(theta * d);
            im -= w * sin(theta * d);
        }
    return re * re + im * im;
}

/* F(k) and dF/dk of the long-range part from the dS/ds table: F(k) = int_0^1 (-dS/ds) cos(k rc s) ds */
static void split_F_from_table(const oracle_esp_in* in, double k, double* F, double* dF)
{
    if (in->split_kind == 1) {
        const double a2 = in->kappa * in->kappa;
        *F = exp(-k * k / (4 * a2));
        *dF = -k / (2 * a2) * (*F);
        return;
    }
    /* 3-point Gauss per segment */
    static const double gx[3] = {0.1127016653792583, 0.5, 0.8872983346207417};
    static const double gw[3] = {0.2777777777777778, 0.4444444444444444, 0.2777777777777778};
    double f = 0, df = 0;
    const double h = 1.0 / in->n_segs;
    for (int s = 0; s < in->n_segs; ++s)
        for (int g = 0; g < 3; ++g) {
            const double t = gx[g];
            const double x = in->r_lo[s] + t / in->dr_inv[s];
            const double w = -gw[g] * h * horner5(in->dscreen, (unsigned)s, t);
            f += w * cos(k * in->rc * x);
            df -= w * in->rc * x * sin(k * in->rc * x);
        }
    *F = f;
    *dF = df;
}

static double split_background(const oracle_esp_in* in)
{
    if (in->split_kind == 1) return 1.0 / (4 * in->kappa * in->kappa);
    static const double gx[3] = {0.1127016653792583, 0.5, 0.8872983346207417};
    static const double gw[3] = {0.2777777777777778, 0.4444444444444444, 0.2777777777777778};
    double m2 = 0;
    const double h = 1.0 / in->n_segs;
    for (int s = 0; s < in->n_segs; ++s)
        for (int g = 0; g < 3; ++g) {
            const double t = gx[g];
            const double x = in->r_lo[s] + t / in->dr_inv[s];
            m2 += -gw[g] * h * horner5(in->dscreen, (unsigned)s, t) * x * x;
        }
    return in->rc * in->rc * m2 / 2.0;
}

typedef struct {
    double key;
    double F, dF;
} fcache_t;

static int cmp_double(const void* a, const void* b)
{
    const double x = *(const double*)a, y = *(const double*)b;
    return (x > y) - (x < y);
}

/* per-target accumulator of the pair correction */
typedef struct {
    const oracle_esp_in* in;
    int i;
    double f[3], e, v[6];
} esp_pair_ctx;

static void pair_accumulate(esp_pair_ctx* c, const double* d, double fdivr, double dE)
{
    for (int a = 0; a < 3; ++a) c->f[a] += fdivr * d[a];
    c->e += 0.5 * dE;
    c->v[0] += 0.5 * d[0] * d[0] * fdivr;
    c->v[1] += 0.5 * d[0] * d[1] * fdivr;
    c->v[2] += 0.5 * d[0] * d[2] * fdivr;
    c->v[3] += 0.5 * d[1] * d[1] * fdivr;
    c->v[4] += 0.5 * d[1] * d[2] * fdivr;
    c->v[5] += 0.5 * d[2] * d[2] * fdivr;
}

/* q_i q_j L(r) for a non-excluded pair within rc (EvaluatorPairPSWF.h:274-295) */
static void esp_pair_term(void* vctx, int j, const double* d, double r2)
{
    esp_pair_ctx* c = (esp_pair_ctx*)vctx;
    const oracle_esp_in* in = c->in;
    const double qq = in->q[c->i] * in->q[j];
    if (qq == 0.0) return;
    if (in->n_ex)
        for (unsigned int k = 0; k < in->n_ex[c->i]; ++k)
            if (in->ex_list[(size_t)c->i * in->ex_pitch + k] == (unsigned)j) return;
    const double r = sqrt(r2);
    double Lr, nd;
    lookup(in->screen, in->dscreen, in->r_lo, in->dr_inv, (unsigned)in->n_segs, in->rc, r, &Lr, &nd, NULL, NULL);
    pair_accumulate(c, d, qq * nd / r, qq * Lr);
}

int oracle_esp(const oracle_esp_in* in, oracle_esp_out* out)
{
    const int N = in->N, P = in->P, D = in->D;
    const int nx = in->n[0], ny = in->n[1], nz = in->n[2];
    const size_t M = (size_t)nx * ny * nz;
    const double* L = in->L;
    const double V = L[0] * L[1] * L[2];
    const double h[3] = {L[0] / nx, L[1] / ny, L[2] / nz};
    const int nm[3] = {nx, ny, nz};
    const obox_t box = {{L[0], L[1], L[2]}, in->tilt[0], in->tilt[1], in->tilt[2]};
    const int tilted = (box.xy != 0.0 || box.xz != 0.0 || box.yz != 0.0);
    /* the pair sum of a tilted box takes the nearest image of every pair */
    if (tilted && in->rc >= box_half_width(&box)) return -1;

    /* ---- assignParticles: spread q_j W(m - g_j) ---- */
    int* i0 = malloc(sizeof(int) * 3 * (size_t)N);
    double* wts = malloc(sizeof(double) * 3 * (size_t)N * P);
    double* dwts = malloc(sizeof(double) * 3 * (size_t)N * P);
    for (int j = 0; j < N; ++j) {
        double uj[3];
        box_fraction3(&box, &in->postype[4 * j], uj); /* makeFraction for a box centred at 0 */
        for (int a = 0; a < 3; ++a) {
            double frac = uj[a];
            frac -= floor(frac);
            const double g = frac * nm[a];
            const double o = ceil(g - 0.5 * P);
            const double t = o - (g - 0.5 * P);
            int oi = (int)o;
            oi = ((oi % nm[a]) + nm[a]) % nm[a];
            i0[3 * j + a] = oi;
            for (int k = 0; k < P; ++k) {
                double dw;
                wts[((size_t)j * 3 + a) * P + k] = window_eval(in->wcoef + (size_t)k * (D + 1), D, t, &dw);
                dwts[((size_t)j * 3 + a) * P + k] = dw;
            }
        }
    }
    double complex* mesh = calloc(M, sizeof(double complex));
    for (int j = 0; j < N; ++j) {
        const double* wx = &wts[((size_t)j * 3 + 0) * P];
        const double* wy = &wts[((size_t)j * 3 + 1) * P];
        const double* wz = &wts[((size_t)j * 3 + 2) * P];
        for (int c = 0; c < P; ++c) {
            const int cz = (i0[3 * j + 2] + c) % nz;
            for (int b = 0; b < P; ++b) {
                const int cy = (i0[3 * j + 1] + b) % ny;
                for (int a = 0; a < P; ++a) {
                    const int cx = (i0[3 * j] + a) % nx;
                    mesh[((size_t)cz * ny + cy) * nx + cx] += in->q[j] * wx[a] * wy[b] * wz[c];
                }
            }
        }
    }
    if (out->mesh)
        for (size_t m = 0; m < M; ++m) out->mesh[m] = creal(mesh[m]);

    /* ---- updateMeshes: forward FFT ---- */
    oracle_fft3d(nx, ny, nz, (double*)mesh, -1);

    /* ---- influence function G(k) = F(|k|) / (k^2 |w(kx)|^2 |w(ky)|^2 |w(kz)|^2) ---- */
    double* w2[3];
    for (int a = 0; a < 3; ++a) {
        w2[a] = malloc(sizeof(double) * nm[a]);
        for (int i = 0; i < nm[a]; ++i) {
            const int ki = (i <= nm[a] / 2) ? i : i - nm[a];
            w2[a][i] = window_fourier2(in->wcoef, P, D, 2 * ORACLE_PI * ki / nm[a]);
        }
    }
    /* distinct k^2 values -> F cache */
    double* k2s = malloc(sizeof(double) * M);
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                const int kx = (x <= nx / 2) ? x : x - nx, ky = (y <= ny / 2) ? y : y - ny, kz = (z <= nz / 2) ? z : z - nz;
                double kv[3];
                box_wavevector(&box, kx, ky, kz, kv);
                k2s[((size_t)z * ny + y) * nx + x] = kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2];
            }
    double* uniq = malloc(sizeof(double) * M);
    memcpy(uniq, k2s, sizeof(double) * M);
    qsort(uniq, M, sizeof(double), cmp_double);
    size_t nu = 0;
    for (size_t m = 0; m < M; ++m)
        if (nu == 0 || uniq[m] != uniq[nu - 1]) uniq[nu++] = uniq[m];
    double* uF = malloc(sizeof(double) * nu);
    double* udF = malloc(sizeof(double) * nu);
#pragma omp parallel for schedule(dynamic, 16)
    for (size_t u = 0; u < nu; ++u) split_F_from_table(in, sqrt(uniq[u]), &uF[u], &udF[u]);

    double Erec = 0, Wrec[6] = {0, 0, 0, 0, 0, 0};
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                const size_t m = ((size_t)z * ny + y) * nx + x;
                const double k2 = k2s[m];
                if (k2 == 0.0 || (in->k_cut > 0.0 && k2 > in->k_cut * in->k_cut)) {
                    mesh[m] = 0;
                    if (out->green) out->green[m] = 0;
                    continue;
                }
                /* binary search in uniq */
                size_t lo = 0, hi = nu - 1;
                while (lo < hi) {
                    size_t mid = (lo + hi) / 2;
                    if (uniq[mid] < k2) lo = mid + 1; else hi = mid;
                }
                double F = uF[lo], dF = udF[lo];
                const double kk = sqrt(k2);
                if (in->alpha > 0.0) {
                    /* screened long-range part F(k) exp(-alpha k) (ESPForceComputeGPU.cu:199) and its k-derivative */
                    const double scr = exp(-in->alpha * kk);
                    dF = (dF - in->alpha * F) * scr;
                    F *= scr;
                }
                const double wprod = w2[0][x] * w2[1][y] * w2[2][z];
                const double G = F / (k2 * wprod);
                /* d/dk (F/k^2) / k  / wprod */
                const double Gv = (dF / (k2 * kk) - 2.0 * F / (k2 * k2)) / wprod;
                if (out->green) out->green[m] = G;
                const double s2 = creal(mesh[m]) * creal(mesh[m]) + cimag(mesh[m]) * cimag(mesh[m]);
                const int kx = (x <= nx / 2) ? x : x - nx, ky = (y <= ny / 2) ? y : y - ny, kz = (z <= nz / 2) ? z : z - nz;
                double k3[3];
                box_wavevector(&box, kx, ky, kz, k3);
                const double e = 0.5 / V * s2;
                Erec += e * G;
                Wrec[0] += e * (G + Gv * k3[0] * k3[0]);
                Wrec[1] += e * (Gv * k3[0] * k3[1]);
                Wrec[2] += e * (Gv * k3[0] * k3[2]);
                Wrec[3] += e * (G + Gv * k3[1] * k3[1]);
                Wrec[4] += e * (Gv * k3[1] * k3[2]);
                Wrec[5] += e * (G + Gv * k3[2] * k3[2]);
                mesh[m] *= G;
            }

    /* ---- inverse FFT: phi(m) = (1/V) sum_k G M^(k) e^{i k x_m} ---- */
    oracle_fft3d(nx, ny, nz, (double*)mesh, +1);
    double* phi = malloc(sizeof(double) * M);
    for (size_t m = 0; m < M; ++m) phi[m] = creal(mesh[m]) / V;
    if (out->phi) memcpy(out->phi, phi, sizeof(double) * M);

    /* ---- interpolateForces: F_j = q_j sum_m grad W(x_m - r_j) phi(m) ---- */
    memset(out->force, 0, sizeof(double) * 4 * (size_t)N);
    for (int c = 0; c < 6; ++c) memset(out->virial + (size_t)c * out->virial_pitch, 0, sizeof(double) * N);
#pragma omp parallel for schedule(static)
    for (int j = 0; j < N; ++j) {
        const double* wx = &wts[((size_t)j * 3 + 0) * P];
        const double* wy = &wts[((size_t)j * 3 + 1) * P];
        const double* wz = &wts[((size_t)j * 3 + 2) * P];
        const double* dx = &dwts[((size_t)j * 3 + 0) * P];
        const double* dy = &dwts[((size_t)j * 3 + 1) * P];
        const double* dz = &dwts[((size_t)j * 3 + 2) * P];
        double ex = 0, ey = 0, ez = 0;
        for (int c = 0; c < P; ++c) {
            const int cz = (i0[3 * j + 2] + c) % nz;
            for (int b = 0; b < P; ++b) {
                const int cy = (i0[3 * j + 1] + b) % ny;
                for (int a = 0; a < P; ++a) {
                    const int cx = (i0[3 * j] + a) % nx;
                    const double p = phi[((size_t)cz * ny + cy) * nx + cx];
                    ex += dx[a] * wy[b] * wz[c] * p;
                    ey += wx[a] * dy[b] * wz[c] * p;
                    ez += wx[a] * wy[b] * dz[c] * p;
                }
            }
        }
        /* gradient along the mesh axes -> Cartesian: d/dr = B^-T d/du */
        const double gx = ex / h[0], gy = ey / h[1], gz = ez / h[2];
        out->force[4 * j + 0] = in->q[j] * gx;
        out->force[4 * j + 1] = in->q[j] * (gy - box.xy * gx);
        out->force[4 * j + 2] = in->q[j] * (gz - box.yz * gy + (box.xy * box.yz - box.xz) * gx);
    }

    /* ---- computeRealSpaceCorrection: pair sum of q_i q_j L(r) over the pairs within rc (cell
     * list; rc <= L/2, so each pair is its nearest image), then fixExclusions for the excluded
     * pairs at any distance ---- */
    cells_t cl;
    if (!tilted) cells_build(&cl, N, in->postype, L, in->rc);
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < N; ++i) {
        esp_pair_ctx c;
        memset(&c, 0, sizeof c);
        c.in = in;
        c.i = i;
        if (!tilted) {
            double ri[3];
            for (int a = 0; a < 3; ++a) {
                double f = in->postype[4 * i + a] / L[a] + 0.5;
                ri[a] = (f - floor(f) - 0.5) * L[a];
            }
            cells_for_neighbours(&cl, in->postype, L, i, ri, in->rc, esp_pair_term, &c);
        } else {
            /* tilted box: every pair at its nearest image (rc below half the smallest width) */
            for (int j = 0; j < N; ++j) {
                if (j == i) continue;
                double d[3] = {in->postype[4 * i] - in->postype[4 * j], in->postype[4 * i + 1] - in->postype[4 * j + 1],
                               in->postype[4 * i + 2] - in->postype[4 * j + 2]};
                min_image_box(d, &box);
                const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
                if (r2 < in->rc * in->rc) esp_pair_term(&c, j, d, r2);
            }
        }
        if (in->n_ex)
            for (unsigned int k = 0; k < in->n_ex[i]; ++k) {
                /* fixExclusions (ESPForceComputeGPU.cu:441-458): remove the mesh's M(r) = 1/(4 pi r) - L(r)
                 * for the excluded pair, at any distance */
                const unsigned int j = in->ex_list[(size_t)i * in->ex_pitch + k];
                double d[3] = {in->postype[4 * i] - in->postype[4 * j], in->postype[4 * i + 1] - in->postype[4 * j + 1],
                               in->postype[4 * i + 2] - in->postype[4 * j + 2]};
                min_image_box(d, &box);
                const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
                if (r2 < 1.0e-20) continue;
                const double r = sqrt(r2);
                const double qq = in->q[i] * in->q[j];
                double Lr = 0, nd = 0;
                if (r <= in->rc)
                    lookup(in->screen, in->dscreen, in->r_lo, in->dr_inv, (unsigned)in->n_segs, in->rc, r, &Lr, &nd, NULL, NULL);
                const double dE = qq * (Lr - ORACLE_INV_4PI / r);
                const double fdivr = qq * (nd - ORACLE_INV_4PI / r2) / r;
                pair_accumulate(&c, d, fdivr, dE);
            }
        out->force[4 * i + 0] += c.f[0];
        out->force[4 * i + 1] += c.f[1];
        out->force[4 * i + 2] += c.f[2];
        out->force[4 * i + 3] += c.e;
        for (int k = 0; k < 6; ++k) out->virial[(size_t)k * out->virial_pitch + i] += c.v[k];
    }
    if (!tilted) cells_free(&cl);

    /* ---- self and neutralising-background terms ---- */
    double q2 = 0, qs = 0;
    for (int j = 0; j < N; ++j) {
        q2 += in->q[j] * in->q[j];
        qs += in->q[j];
    }
    /* lim_{r->0} (1 - S)/r = -dS/dr(0) */
    double self_coeff;
    if (in->split_kind == 1)
        self_coeff = 2.0 * in->kappa / sqrt(ORACLE_PI);
    else
        self_coeff = -horner5(in->dscreen, 0, 0.0) / in->rc;
    const double Eself = -0.5 * ORACLE_INV_4PI * self_coeff * q2;
    const double Ebg = -qs * qs / (2.0 * V) * split_background(in);
    out->ext_energy = Erec + Eself + Ebg;
    for (int c = 0; c < 6; ++c) out->ext_virial[c] = Wrec[c];
    out->ext_virial[0] += Ebg;
    out->ext_virial[3] += Ebg;
    out->ext_virial[5] += Ebg;

    free(i0);
    free(wts);
    free(dwts);
    free(mesh);
    for (int a = 0; a < 3; ++a) free(w2[a]);
    free(k2s);
    free(uniq);
    free(uF);
    free(udF);
    free(phi);
    return 0;
}

/* Window / influence pieces exposed for unit tests */
double oracle_window_fourier2(const double* wc, int P, int D, double theta)
{
    return window_fourier2(wc, P, D, theta);
}

void oracle_window_eval(const double* wc, int P, int D, double t, double* w, double* dw)
{
    for (int k = 0; k < P; ++k) w[k] = window_eval(wc + (size_t)k * (D + 1), D, t, &dw[k]);
}

void oracle_split_F(const oracle_esp_in* in, double k, double* F, double* dF)
{
    split_F_from_table(in, k, F, dF);
}
