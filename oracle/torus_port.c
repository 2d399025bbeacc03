/* Test and baseline infrastructure (never linked into the engine): CGGI functional
 * bootstrapping over the 32-bit torus in C, the computation oracle/torus.py states
 * (and tests/test_torus_oracle.py pins), used as the CPU baseline of that bootstrap and
 * checked against oracle/torus.py bit for bit.  There is no reference counterpart (the
 * reference's CGGI16 is the RNS variant, modules/fhe/python/src/vfhe/fhe/cggi16.py:46-140);
 * the structure follows it: rotation amounts (cggi16.py:60-62,77, q = 2^32, ties to even),
 * the test vector rotated by -rot_b, one CMUX per non-zero amount, extraction of
 * coefficient 0.  External products are exact integers: digits |d| <= 2^(B-1) times
 * centred 32-bit key rows through a double FFT whose error stays far below 1/2.
 *
 * Negacyclic transform: z_j = (p_j + i p_(j+M)) zeta^j (zeta = e^(i pi / N), M = N/2),
 * then an M-point DFT with w = e^(-2 pi i / M) (iterative radix-2, natural order in and
 * out); the inverse undoes both.  Samples are spread over POSIX threads.
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    int N, M, logm, n, ell, B;
    uint32_t offset, dec_off;
    double *twr, *twi;      /* w^k, k < M/2 */
    double *zr, *zi;        /* zeta^j, j < M */
    int *brv;               /* bit reversal of log2 M bits */
    double *kr, *ki;        /* key rows transformed: [(i * 2 ell + row) * 2 + comp][M] */
    const uint32_t *lwe, *tv;
    uint32_t *out;
    uint64_t count, next;
    pthread_mutex_t mu;
} tp_ctx;

static void fft(const tp_ctx *C, double *re, double *im, int inverse)
{
    const int M = C->M;
    for (int k = 0; k < M; k++) {
        const int r = C->brv[k];
        if (r > k) {
            double t = re[k]; re[k] = re[r]; re[r] = t;
            t = im[k]; im[k] = im[r]; im[r] = t;
        }
    }
    for (int len = 2; len <= M; len <<= 1) {
        const int half = len >> 1, step = M / len;
        for (int s = 0; s < M; s += len)
            for (int j = 0; j < half; j++) {
                const double wr = C->twr[j * step], wi = inverse ? -C->twi[j * step] : C->twi[j * step];
                const int a = s + j, b = a + half;
                const double xr = re[b] * wr - im[b] * wi, xi = re[b] * wi + im[b] * wr;
                re[b] = re[a] - xr;
                im[b] = im[a] - xi;
                re[a] += xr;
                im[a] += xi;
            }
    }
}

/* forward transform of a polynomial of N signed values */
static void fwd(const tp_ctx *C, const double *p, double *re, double *im)
{
    const int M = C->M;
    for (int j = 0; j < M; j++) {
        const double a = p[j], b = p[j + M];
        re[j] = a * C->zr[j] - b * C->zi[j];
        im[j] = a * C->zi[j] + b * C->zr[j];
    }
    fft(C, re, im, 0);
}

/* inverse transform, rounded to integers mod 2^32 */
static void inv_round(const tp_ctx *C, double *re, double *im, uint32_t *p)
{
    const int M = C->M;
    fft(C, re, im, 1);
    for (int j = 0; j < M; j++) {
        const double a = re[j] / M, b = im[j] / M;
        const double cr = a * C->zr[j] + b * C->zi[j], ci = b * C->zr[j] - a * C->zi[j];
        p[j] = (uint32_t)(int64_t)llround(cr);
        p[j + M] = (uint32_t)(int64_t)llround(ci);
    }
}

/* round(x 2N / 2^32) mod 2N with ties to even */
static uint32_t mod_switch(uint64_t x, int N)
{
    const int shift = 32 - (int)log2(2.0 * N);
    const uint64_t q = x >> shift, rem = x & ((1ull << shift) - 1), half = 1ull << (shift - 1);
    const uint64_t r = q + ((rem > half || (rem == half && (q & 1))) ? 1 : 0);
    return (uint32_t)(r & (uint64_t)(2 * N - 1));
}

/* coefficient j of X^a p */
static uint32_t rot(const uint32_t *p, int N, int j, uint32_t a)
{
    const uint32_t t = ((uint32_t)j + 2u * (uint32_t)N - a) & (uint32_t)(2 * N - 1);
    return t < (uint32_t)N ? p[t] : 0u - p[t - N];
}

static void bootstrap_one(const tp_ctx *C, uint64_t k, double *ws)
{
    const int N = C->N, M = C->M, n = C->n, ell = C->ell, B = C->B, rows = 2 * ell;
    const uint32_t *in = C->lwe + k * (uint64_t)(n + 1);
    uint32_t *acc = (uint32_t *)ws;                        /* 2 N */
    uint32_t *diff = acc + 2 * N;                          /* 2 N */
    uint32_t *res = diff + 2 * N;                          /* N */
    double *dig = (double *)(res + N + (N & 1));           /* N */
    double *fr = dig + N, *fi = fr + M;                    /* digit transform */
    double *pr = fi + M, *pi = pr + 2 * M;                 /* 2 output accumulators */
    const uint32_t rot_b = mod_switch((uint64_t)in[n] + C->offset, N);
    const uint32_t rot_tv = (uint32_t)(2 * N - rot_b) & (uint32_t)(2 * N - 1);
    for (int j = 0; j < N; j++) {
        acc[j] = 0;
        acc[N + j] = rot(C->tv, N, j, rot_tv);
    }
    const uint32_t bmask = (1u << B) - 1, bhalf = 1u << (B - 1);
    for (int i = 0; i < n; i++) {
        const uint32_t a = mod_switch(in[i], N);
        if (a == 0) continue;
        for (int c = 0; c < 2; c++)
            for (int j = 0; j < N; j++) diff[c * N + j] = rot(acc + c * N, N, j, a) - acc[c * N + j] + C->dec_off;
        memset(pr, 0, sizeof(double) * 4 * M);
        for (int c = 0; c < 2; c++)
            for (int l = 0; l < ell; l++) {
                const int sh = 32 - B * (l + 1);
                for (int j = 0; j < N; j++) dig[j] = (double)(int)(((diff[c * N + j] >> sh) & bmask) - bhalf);
                fwd(C, dig, fr, fi);
                const size_t row = ((size_t)i * rows + (size_t)c * ell + l) * 2;
                for (int o = 0; o < 2; o++) {
                    const double *kr = C->kr + (row + o) * M, *ki = C->ki + (row + o) * M;
                    double *qr = pr + o * M, *qi = pi + o * M;
                    for (int m = 0; m < M; m++) {
                        qr[m] += fr[m] * kr[m] - fi[m] * ki[m];
                        qi[m] += fr[m] * ki[m] + fi[m] * kr[m];
                    }
                }
            }
        for (int o = 0; o < 2; o++) {
            inv_round(C, pr + o * M, pi + o * M, res);
            for (int j = 0; j < N; j++) acc[o * N + j] += res[j];
        }
    }
    uint32_t *o = C->out + k * (uint64_t)(N + 1);
    o[0] = acc[0];
    for (int j = 1; j < N; j++) o[j] = 0u - acc[N - j];
    o[N] = acc[N];
}

static void *worker(void *arg)
{
    tp_ctx *C = (tp_ctx *)arg;
    const size_t words = 5 * (size_t)C->N + 2 + 4 * (size_t)C->N;
    double *ws = (double *)malloc(words * sizeof(double));
    if (!ws) return NULL;
    for (;;) {
        pthread_mutex_lock(&C->mu);
        const uint64_t k = C->next++;
        pthread_mutex_unlock(&C->mu);
        if (k >= C->count) break;
        bootstrap_one(C, k, ws);
    }
    free(ws);
    return NULL;
}

/* The key rows (bsk: n x 2 ell x 2 x N torus words) transformed once: a context for
 * torus_port_run, freed with torus_port_free.  NULL on bad sizes. */
void *torus_port_new(const uint32_t *bsk, int n, int N, int ell, int B)
{
    if (N < 4 || (N & (N - 1)) || n < 1 || ell < 1 || B < 1 || ell * B > 32) return NULL;
    tp_ctx *C = calloc(1, sizeof(tp_ctx));
    if (!C) return NULL;
    C->N = N;
    C->M = N / 2;
    C->n = n;
    C->ell = ell;
    C->B = B;
    uint32_t off = 0;
    for (int l = 1; l <= ell; l++) off += 1u << (31 - B * (l - 1));
    if (B * ell < 32) off += 1u << (32 - B * ell - 1);
    C->dec_off = off;
    const int M = C->M;
    int logm = 0;
    while ((1 << logm) < M) logm++;
    C->logm = logm;
    C->twr = malloc(sizeof(double) * M);
    C->twi = malloc(sizeof(double) * M);
    C->zr = malloc(sizeof(double) * M);
    C->zi = malloc(sizeof(double) * M);
    C->brv = malloc(sizeof(int) * M);
    const size_t krows = (size_t)n * 2 * ell * 2;
    C->kr = malloc(sizeof(double) * krows * M);
    C->ki = malloc(sizeof(double) * krows * M);
    double *tmp = malloc(sizeof(double) * N);
    if (!C->twr || !C->twi || !C->zr || !C->zi || !C->brv || !C->kr || !C->ki || !tmp) abort();
    const long double pi = 3.141592653589793238462643383279502884L;
    for (int k = 0; k < M; k++) {
        C->twr[k] = (double)cosl(-2 * pi * k / M);
        C->twi[k] = (double)sinl(-2 * pi * k / M);
        C->zr[k] = (double)cosl(pi * k / N);
        C->zi[k] = (double)sinl(pi * k / N);
        int r = 0;
        for (int b = 0; b < logm; b++) r |= ((k >> b) & 1) << (logm - 1 - b);
        C->brv[k] = r;
    }
    for (size_t r = 0; r < krows; r++) {
        for (int j = 0; j < N; j++) tmp[j] = (double)(int32_t)bsk[r * N + j];
        fwd(C, tmp, C->kr + r * M, C->ki + r * M);
    }
    free(tmp);
    return C;
}

void torus_port_free(void *ctx)
{
    tp_ctx *C = (tp_ctx *)ctx;
    if (!C) return;
    free(C->twr); free(C->twi); free(C->zr); free(C->zi); free(C->brv); free(C->kr); free(C->ki);
    free(C);
}

/* lwe: count x (n + 1); tv: N; out: count x (N + 1).  Returns 0, or -1 on bad arguments. */
int torus_port_run(void *ctx, const uint32_t *lwe, uint64_t count, const uint32_t *tv, uint64_t torus_base,
                   uint32_t *out, int threads)
{
    tp_ctx *C = (tp_ctx *)ctx;
    if (!C || !torus_base || threads < 1) return -1;
    C->offset = (uint32_t)((1ull << 32) / (4 * torus_base));
    C->lwe = lwe;
    C->tv = tv;
    C->out = out;
    C->count = count;
    C->next = 0;
    pthread_mutex_init(&C->mu, NULL);
    pthread_t th[256];
    const int nt = threads > 256 ? 256 : threads;
    for (int t = 0; t < nt; t++) pthread_create(&th[t], NULL, worker, C);
    for (int t = 0; t < nt; t++) pthread_join(th[t], NULL);
    pthread_mutex_destroy(&C->mu);
    return 0;
}
