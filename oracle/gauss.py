"""Float64 oracle of the separable Gaussian image model and its gradients.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py): imported by tests/ and
bench.py's cpu_baseline leg, never by the product.

Model (pixel units; pixel i covers [i, i + 1)): emitter (x, y, z, I) adds
    I * exp(-|z| * inv_depth) * G(i; x) * G(j; y)
to pixel (i, j) of its frame for i in [floor(x) - R, floor(x) + R] (same for
j), with G(i; x) the unit Gaussian of width sigma integrated over the pixel:
    G(i; x) = (erf((i + 1 - x) / (sqrt(2) sigma)) - erf((i - x) / (sqrt(2) sigma))) / 2.
It is the Gaussian radial profile of src/scopyon/_epifm.py:128-134 integrated
exactly over each pixel, weighted by the TIRF amplitude of _epifm.py:1289.
The image is background + the sum; the loss is the Poisson negative
log-likelihood sum(mu - k log mu) per frame.

Pins (tests/test_oracle_gauss.py): G against numerical quadrature of the
normal density, the telescoping sum of G over a window, and the analytic
gradients against float64 central finite differences of ``render``.
"""
import numpy
from scipy.special import erf, erfc

SQRT2 = numpy.sqrt(2.0)


def g_pixel(i, x, sigma):
    """Pixel-integrated unit Gaussian, complementary form in the tails."""
    i = numpy.asarray(i, dtype=float)
    t0 = (i - x) / (SQRT2 * sigma)
    t1 = (i + 1.0 - x) / (SQRT2 * sigma)
    out = 0.5 * (erf(t1) - erf(t0))
    right = t0 >= 0
    left = t1 <= 0
    out = numpy.where(right, 0.5 * (erfc(t0) - erfc(t1)), out)
    out = numpy.where(left, 0.5 * (erfc(-t1) - erfc(-t0)), out)
    return out


def normal_pdf(t, sigma):
    return numpy.exp(-0.5 * (t / sigma) ** 2) / (numpy.sqrt(2.0 * numpy.pi) * sigma)


def render(shapes, x, y, z, I, frame_id, sigma, inv_depth, background, radius):
    """List of float64 frames (mu) for emitters with the given frame ids."""
    frames = [numpy.full(s, float(background)) for s in shapes]
    W = 2 * radius + 1
    k = numpy.arange(W)
    for e in range(len(x)):
        f = int(frame_id[e])
        nw, nh = shapes[f]
        lu = int(numpy.floor(x[e])) - radius
        lv = int(numpy.floor(y[e])) - radius
        gx = g_pixel(lu + k, x[e], sigma)
        gy = g_pixel(lv + k, y[e], sigma)
        amp = I[e] * numpy.exp(-abs(z[e]) * inv_depth)
        patch = amp * numpy.outer(gx, gy)
        r0, r1 = max(lu, 0), min(lu + W, nw)
        c0, c1 = max(lv, 0), min(lv + W, nh)
        if r1 > r0 and c1 > c0:
            frames[f][r0:r1, c0:c1] += patch[r0 - lu:r1 - lu, c0 - lv:c1 - lv]
    return frames


def poisson_nll(frames, targets):
    """Per-frame sum of mu - k log(mu)."""
    return numpy.array([float(numpy.sum(m - k * numpy.log(m))) for m, k in zip(frames, targets)])


def gradients(shapes, x, y, z, I, frame_id, sigma, inv_depth, background, radius, targets):
    """Analytic dL/d(x, y, z, I) per emitter and dL/dsigma for
    L = sum over frames of poisson_nll."""
    mu = render(shapes, x, y, z, I, frame_id, sigma, inv_depth, background, radius)
    g_img = [1.0 - k / m for m, k in zip(mu, targets)]
    n = len(x)
    gx, gy, gz, gI = (numpy.zeros(n) for _ in range(4))
    gsig = 0.0
    W = 2 * radius + 1
    k = numpy.arange(W)
    for e in range(n):
        f = int(frame_id[e])
        nw, nh = shapes[f]
        lu = int(numpy.floor(x[e])) - radius
        lv = int(numpy.floor(y[e])) - radius
        iu, jv = lu + k, lv + k
        Gx, Gy = g_pixel(iu, x[e], sigma), g_pixel(jv, y[e], sigma)
        a0, a1 = iu - x[e], iu + 1.0 - x[e]
        b0, b1 = jv - y[e], jv + 1.0 - y[e]
        dGx = normal_pdf(a0, sigma) - normal_pdf(a1, sigma)
        dGy = normal_pdf(b0, sigma) - normal_pdf(b1, sigma)
        sGx = -(a1 * normal_pdf(a1, sigma) - a0 * normal_pdf(a0, sigma)) / sigma
        sGy = -(b1 * normal_pdf(b1, sigma) - b0 * normal_pdf(b0, sigma)) / sigma
        g = numpy.zeros((W, W))
        r0, r1 = max(lu, 0), min(lu + W, nw)
        c0, c1 = max(lv, 0), min(lv + W, nh)
        if r1 > r0 and c1 > c0:
            g[r0 - lu:r1 - lu, c0 - lv:c1 - lv] = g_img[f][r0:r1, c0:c1]
        a = numpy.exp(-abs(z[e]) * inv_depth)
        S_I = Gx @ g @ Gy
        gI[e] = a * S_I
        gx[e] = I[e] * a * (dGx @ g @ Gy)
        gy[e] = I[e] * a * (Gx @ g @ dGy)
        gz[e] = -numpy.sign(z[e]) * inv_depth * I[e] * a * S_I
        gsig += I[e] * a * (sGx @ g @ Gy + Gx @ g @ sGy)
    return gx, gy, gz, gI, gsig


def loss(shapes, x, y, z, I, frame_id, sigma, inv_depth, background, radius, targets):
    mu = render(shapes, x, y, z, I, frame_id, sigma, inv_depth, background, radius)
    return float(numpy.sum(poisson_nll(mu, targets)))
