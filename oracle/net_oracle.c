/*
 * net_oracle.c -- FP64 restatement of the defocus-proposal network (forward,
 * loss, backward), the checker of the HIP trainer's kernels.
 *
 * TEST INFRASTRUCTURE ONLY.  tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product library never does.
 *
 * The network has no counterpart in the reference.  Definition (the same
 * parameter layout as the HIP trainer, one flat vector):
 *   input   radial bins [16, 256) z-scored per spectrum (population
 *           variance), bins [0, 16) zero
 *   conv1   1 -> 64, 7 taps, zero padding 3, +bias, ReLU, mean-pool 2
 *   conv2-4 64 -> 64, 7 taps (weight [co][ci * 7 + t]), same
 *   fc1     1024 -> 128 (+bias, ReLU), input index c * 16 + l
 *   fc2     128 -> 4
 *   loss    mean over the batch of
 *           100 (o0 - ln(df / 1e4))^2 + 0.25 (o1 - ast / 1e3)^2
 *           + 0.05 (1 - cos 2 (o2 - az)) + 0.5 w BCE(sigmoid(o3), c),
 *           c = 1 if target[3] > 0, 0 if < 0; w = 0 when target[3] == 0
 *           (no confidence label), else 1
 *
 * Parity pin: the gradient of orc_net_loss_grad is checked against central
 * finite differences of orc_net_loss_grad's own loss (tests/test_net.py),
 * and the loss terms against closed forms on hand-set outputs.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NB 256
#define NB0 16
#define CH 64
#define TAPS 7
#define HID 128
#define NOUT 4
#define L4 16
#define FLAT (CH * L4)
#define COL (CH * TAPS)

#define OW1 0L
#define OB1 (OW1 + CH * TAPS)
#define OW2 (OB1 + CH)
#define OB2 (OW2 + (long)CH * COL)
#define OW3 (OB2 + CH)
#define OB3 (OW3 + (long)CH * COL)
#define OW4 (OB3 + CH)
#define OB4 (OW4 + (long)CH * COL)
#define OW5 (OB4 + CH)
#define OB5 (OW5 + (long)HID * FLAT)
#define OW6 (OB5 + HID)
#define OB6 (OW6 + (long)NOUT * HID)
#define NPARAM (OB6 + NOUT)

long orc_net_param_count(void) { return NPARAM; }

/* per-sample activations */
typedef struct {
	double x0[NB];
	double y1[CH][NB], a1[CH][NB / 2];
	double y2[CH][NB / 2], a2[CH][NB / 4];
	double y3[CH][NB / 4], a3[CH][NB / 8];
	double y4[CH][NB / 8], a4[CH][NB / 16];
	double f[FLAT], h[HID], o[NOUT];
} act_t;

static void input(const float* r, double* x0)
{
	double s = 0.0, s2 = 0.0;
	for (int t = NB0; t < NB; ++t) { s += r[t]; s2 += (double)r[t] * r[t]; }
	const double cnt = NB - NB0, mean = s / cnt;
	double var = s2 / cnt - mean * mean;
	if (var < 1e-24) var = 1e-24;
	for (int t = 0; t < NB; ++t) x0[t] = t >= NB0 ? ((double)r[t] - mean) / sqrt(var) : 0.0;
}

/* conv (cin -> CH) + bias + ReLU + pool; in [cin][L], w [co][cin * 7 + t] */
static void conv(const double* p, long wo, long bo, int cin, int L, const double* in, double* y, double* a)
{
	for (int c = 0; c < CH; ++c) {
		for (int l = 0; l < L; ++l) {
			double z = p[bo + c];
			for (int ci = 0; ci < cin; ++ci)
				for (int t = 0; t < TAPS; ++t) {
					const int j = l + t - 3;
					if (j >= 0 && j < L) z += p[wo + (long)c * cin * TAPS + ci * TAPS + t] * in[ci * L + j];
				}
			y[c * L + l] = z > 0.0 ? z : 0.0;
		}
		for (int q = 0; q < L / 2; ++q) a[c * (L / 2) + q] = 0.5 * (y[c * L + 2 * q] + y[c * L + 2 * q + 1]);
	}
}

static void forward(const double* p, const float* r, act_t* A)
{
	input(r, A->x0);
	conv(p, OW1, OB1, 1, NB, A->x0, &A->y1[0][0], &A->a1[0][0]);
	conv(p, OW2, OB2, CH, NB / 2, &A->a1[0][0], &A->y2[0][0], &A->a2[0][0]);
	conv(p, OW3, OB3, CH, NB / 4, &A->a2[0][0], &A->y3[0][0], &A->a3[0][0]);
	conv(p, OW4, OB4, CH, NB / 8, &A->a3[0][0], &A->y4[0][0], &A->a4[0][0]);
	for (int c = 0; c < CH; ++c)
		for (int l = 0; l < L4; ++l) A->f[c * L4 + l] = A->a4[c][l];
	for (int j = 0; j < HID; ++j) {
		double z = p[OB5 + j];
		for (int i = 0; i < FLAT; ++i) z += p[OW5 + (long)j * FLAT + i] * A->f[i];
		A->h[j] = z > 0.0 ? z : 0.0;
	}
	for (int k = 0; k < NOUT; ++k) {
		double z = p[OB6 + k];
		for (int j = 0; j < HID; ++j) z += p[OW6 + k * HID + j] * A->h[j];
		A->o[k] = z;
	}
}

void orc_net_forward(const double* p, const float* radial, int ld, int n, double* out)
{
	act_t* A = (act_t*)malloc(sizeof(act_t));
	for (int b = 0; b < n; ++b) {
		forward(p, radial + (long)b * ld, A);
		for (int k = 0; k < NOUT; ++k) out[b * NOUT + k] = A->o[k];
	}
	free(A);
}

/* loss of one sample and dL/do (not yet divided by the batch size) */
static double loss1(const double* o, const float* tg, double* d)
{
	const double y0 = log(tg[0] * 1e-4), y1 = tg[1] * 1e-3, az = tg[2] * (M_PI / 180.0);
	const double e0 = o[0] - y0, e1 = o[1] - y1, d2 = 2.0 * (o[2] - az);
	/* label sign: > 0 agrees (c = 1), < 0 disagrees (c = 0), 0 no label (w = 0) */
	const double c = tg[3] > 0.0f ? 1.0 : 0.0;
	const double w = tg[3] != 0.0f ? 1.0 : 0.0;
	const double x = o[3];
	const double sp_pos = (x > 0 ? x : 0.0) + log1p(exp(-fabs(x)));
	const double sp_neg = sp_pos - x;
	const double bce = c * sp_neg + (1.0 - c) * sp_pos;
	const double p = 1.0 / (1.0 + exp(-x));
	d[0] = 200.0 * e0;
	d[1] = 0.5 * e1;
	d[2] = 0.1 * sin(d2);
	d[3] = 0.5 * w * (p - c);
	return 100.0 * e0 * e0 + 0.25 * e1 * e1 + 0.05 * (1.0 - cos(d2)) + 0.5 * w * bce;
}

/* backward of conv (cin -> CH, length L): dA_out [CH][L/2] -> grads and dIn [cin][L] */
static void conv_back(const double* p, long wo, long bo, int cin, int L, const double* in, const double* y,
	const double* dA, double* g, double* din)
{
	double* dz = (double*)malloc(sizeof(double) * CH * L);
	for (int c = 0; c < CH; ++c)
		for (int l = 0; l < L; ++l) dz[c * L + l] = y[c * L + l] > 0.0 ? 0.5 * dA[c * (L / 2) + l / 2] : 0.0;
	if (din) memset(din, 0, sizeof(double) * cin * L);
	for (int c = 0; c < CH; ++c)
		for (int l = 0; l < L; ++l) {
			const double d = dz[c * L + l];
			if (d == 0.0) continue;
			g[bo + c] += d;
			for (int ci = 0; ci < cin; ++ci)
				for (int t = 0; t < TAPS; ++t) {
					const int j = l + t - 3;
					if (j < 0 || j >= L) continue;
					const long wi = wo + (long)c * cin * TAPS + ci * TAPS + t;
					g[wi] += d * in[ci * L + j];
					if (din) din[ci * L + j] += d * p[wi];
				}
		}
	free(dz);
}

/* mean loss over n samples; grad (NPARAM doubles) gets its gradient */
double orc_net_loss_grad(const double* p, const float* radial, int ld, const float* target, int n, double* grad)
{
	act_t* A = (act_t*)malloc(sizeof(act_t));
	double* dA = (double*)malloc(sizeof(double) * CH * NB);
	double* dB = (double*)malloc(sizeof(double) * CH * NB);
	if (grad) memset(grad, 0, sizeof(double) * NPARAM);
	double L = 0.0;
	for (int b = 0; b < n; ++b) {
		forward(p, radial + (long)b * ld, A);
		double dO[NOUT];
		L += loss1(A->o, target + 4L * b, dO);
		if (!grad) continue;
		for (int k = 0; k < NOUT; ++k) dO[k] /= n;
		double dh[HID], df[FLAT];
		for (int k = 0; k < NOUT; ++k) {
			grad[OB6 + k] += dO[k];
			for (int j = 0; j < HID; ++j) grad[OW6 + k * HID + j] += dO[k] * A->h[j];
		}
		for (int j = 0; j < HID; ++j) {
			double s = 0.0;
			for (int k = 0; k < NOUT; ++k) s += p[OW6 + k * HID + j] * dO[k];
			dh[j] = A->h[j] > 0.0 ? s : 0.0;
		}
		memset(df, 0, sizeof df);
		for (int j = 0; j < HID; ++j) {
			if (dh[j] == 0.0) continue;
			grad[OB5 + j] += dh[j];
			for (int i = 0; i < FLAT; ++i) {
				grad[OW5 + (long)j * FLAT + i] += dh[j] * A->f[i];
				df[i] += p[OW5 + (long)j * FLAT + i] * dh[j];
			}
		}
		/* unflatten into dA [CH][16] */
		for (int c = 0; c < CH; ++c)
			for (int l = 0; l < L4; ++l) dA[c * L4 + l] = df[c * L4 + l];
		conv_back(p, OW4, OB4, CH, NB / 8, &A->a3[0][0], &A->y4[0][0], dA, grad, dB);
		conv_back(p, OW3, OB3, CH, NB / 4, &A->a2[0][0], &A->y3[0][0], dB, grad, dA);
		conv_back(p, OW2, OB2, CH, NB / 2, &A->a1[0][0], &A->y2[0][0], dA, grad, dB);
		conv_back(p, OW1, OB1, 1, NB, A->x0, &A->y1[0][0], dB, grad, NULL);
	}
	free(A);
	free(dA);
	free(dB);
	return L / n;
}
