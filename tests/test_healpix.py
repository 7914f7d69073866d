"""Pins of the HEALPix NESTED implementation (frt_healpix.h) that stands in for
the un-vendored HEALPix C++ 3.40 library on both the GPU and the CPU checker.

Checked against what Gorski et al. 2005 fixes: the 12 base-pixel centres, the
ring latitudes, the nested hierarchy (children 4p..4p+3 inside the parent),
equal pixel areas, and pix2vec/vec2pix being inverse.
"""
import math

import numpy as np
import pytest

from oracle import pyoracle as P


def unit_vectors(n, seed):
    rng = np.random.default_rng(seed)
    v = rng.standard_normal((n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def test_base_pixel_centres_closed_form():
    # faces 0-3: z = 2/3, phi = pi/4 + k pi/2; faces 4-7: z = 0, phi = k pi/2;
    # faces 8-11: z = -2/3, phi = pi/4 + k pi/2
    for f in range(12):
        v = P.pix2vec(0, f)
        k = f % 4
        if f < 4:
            z, phi = 2.0 / 3.0, math.pi / 4 + k * math.pi / 2
        elif f < 8:
            z, phi = 0.0, k * math.pi / 2
        else:
            z, phi = -2.0 / 3.0, math.pi / 4 + k * math.pi / 2
        st = math.sqrt(1 - z * z)
        assert np.allclose(v, [st * math.cos(phi), st * math.sin(phi), z], atol=1e-15), f


@pytest.mark.parametrize("order", [1, 2, 3])
def test_ring_latitudes(order):
    # ring i (1..4N-1): z = 1 - i^2/(3N^2) (i < N), 4/3 - 2i/(3N) (N <= i <= 3N),
    # mirrored in the south; ring i holds 4i (polar) or 4N (equatorial) pixels
    N = 1 << order
    zs = np.array([P.pix2vec(order, p)[2] for p in range(12 * N * N)])
    expect = []
    for i in range(1, 4 * N):
        if i < N:
            z, cnt = 1 - i * i / (3 * N * N), 4 * i
        elif i <= 3 * N:
            z, cnt = 4 / 3 - 2 * i / (3 * N), 4 * N
        else:
            j = 4 * N - i
            z, cnt = -(1 - j * j / (3 * N * N)), 4 * j
        expect.append((z, cnt))
    for z, cnt in expect:
        assert np.sum(np.abs(zs - z) < 1e-12) == cnt


@pytest.mark.parametrize("order", [0, 1, 2, 3, 4, 5])
def test_roundtrip_all_pixels(order):
    npix = 12 << (2 * order)
    for p in range(npix):
        v = P.pix2vec(order, p)
        assert abs(np.linalg.norm(v) - 1.0) < 1e-15
        assert P.vec2pix(order, v) == p


@pytest.mark.parametrize("order", [8, 10, 12])
def test_roundtrip_sampled_high_order(order):
    rng = np.random.default_rng(order)
    npix = 12 << (2 * order)
    for p in rng.integers(0, npix, 3000):
        assert P.vec2pix(order, P.pix2vec(order, int(p))) == int(p)


def test_nesting_consistency():
    # vec2pix at order l, shifted by 2 bits, is vec2pix at order l-1
    for v in unit_vectors(4000, 3):
        pix = [P.vec2pix(l, v) for l in range(13)]
        for l in range(1, 13):
            assert pix[l] >> 2 == pix[l - 1]


def test_vec2pix_ignores_length():
    for v in unit_vectors(500, 4):
        for l in (0, 3, 7):
            assert P.vec2pix(l, v) == P.vec2pix(l, 3.7 * v)


def test_equal_area():
    # uniform directions land in each of the 192 order-2 pixels equally often
    v = unit_vectors(200000, 5)
    counts = np.bincount([P.vec2pix(2, x) for x in v], minlength=192)
    expect = len(v) / 192
    chi2 = float(((counts - expect) ** 2 / expect).sum())
    # 191 degrees of freedom: mean 191, sd ~19.5; a wrong area would give >> 300
    assert chi2 < 300, chi2


def test_children_inside_parent():
    # the centre of child 4p+c lies within the parent pixel
    for order in range(0, 6):
        for p in range(0, 12 << (2 * order), 7):
            for c in range(4):
                v = P.pix2vec(order + 1, 4 * p + c)
                assert P.vec2pix(order, v) == p


# Gorski et al. 2005: face f of the 12 base pixels sits on ring jrll[f] N and
# starts at in-ring position jpll[f] (in units of half a pixel); inside a
# face the NESTED index interleaves x (even bits) and y (odd bits), with x
# toward the north-east and y toward the north-west.
JRLL = [2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4]
JPLL = [1, 3, 5, 7, 0, 2, 4, 6, 1, 3, 5, 7]


def _deinterleave(v):
    x = y = 0
    b = 0
    while v:
        x |= (v & 1) << b
        y |= ((v >> 1) & 1) << b
        v >>= 2
        b += 1
    return x, y


def _ring_position(order, p):
    """(ring index i, in-ring index j from 1) of NESTED pixel p."""
    N = 1 << order
    f, sub = p >> (2 * order), p & (N * N - 1)
    ix, iy = _deinterleave(sub)
    i = JRLL[f] * N - ix - iy - 1
    nr = i if i < N else (4 * N - i if i > 3 * N else N)
    shift = ((i - N) & 1) if N <= i <= 3 * N else 0
    j = (JPLL[f] * nr + ix - iy + 1 + shift) // 2
    j = j - 4 * N if j > 4 * N else (j + 4 * N if j < 1 else j)
    return i, j


@pytest.mark.parametrize("order", [1, 2, 3])
def test_longitude_of_every_nested_pixel(order):
    """phi of each NESTED pixel centre from its ring and in-ring index:
    polar ring i holds 4i pixels at phi = (pi/2i)(j - 1/2); equatorial ring i
    holds 4N at phi = (pi/2N)(j - (1 + (i - N) mod 2)/2), every other ring
    shifted by half a pixel (Gorski et al. 2005, eqs. 8-9, with j counted
    from 1 as in the HEALPix library).  A swap of the x and y bits inside a
    face, or a shifted ring, moves phi."""
    N = 1 << order
    seen = set()
    for p in range(12 * N * N):
        i, j = _ring_position(order, p)
        seen.add((i, j))
        if i < N:
            phi = math.pi / (2 * i) * (j - 0.5)
        elif i > 3 * N:
            phi = math.pi / (2 * (4 * N - i)) * (j - 0.5)
        else:
            phi = math.pi / (2 * N) * (j - (1 + (i - N) % 2) / 2)
        v = P.pix2vec(order, p)
        got = math.atan2(v[1], v[0]) % (2 * math.pi)
        d = abs(got - phi % (2 * math.pi))
        assert min(d, 2 * math.pi - d) < 1e-12, (p, i, j, got, phi)
    assert len(seen) == 12 * N * N  # the ring positions are a bijection


def test_children_orientation():
    """Inside each base pixel the four order-1 children run south (0), east
    (1, the x direction), west (2, y) and north (3)."""
    for f in range(12):
        z = [P.pix2vec(1, 4 * f + c)[2] for c in range(4)]
        phi = [math.atan2(P.pix2vec(1, 4 * f + c)[1], P.pix2vec(1, 4 * f + c)[0]) for c in range(4)]
        assert z[0] < z[1] and z[0] < z[2] and z[3] > z[1] and z[3] > z[2]
        assert abs(z[1] - z[2]) < 1e-12
        dphi = (phi[1] - phi[2] + math.pi) % (2 * math.pi) - math.pi
        assert dphi > 0, f
