"""A CPU model of the solver's substitution by wavefront affine scans
(bolt.jl_amd/csrc/bolt_hip.hip: wave_subst, pass_scan, affine_scan), with the
kernel's DPP controls and row masks (row_shr 1, 2, 4, 8; row_bcast 15 to rows
1 and 3; row_bcast 31 to rows 2 and 3; wave_shr 1; invalid sources keep the
old value, the identity map), its SR rows per lane, its passes of 64·SR rows,
its masking of rows past n in the backward pass, and the two second-band
corrections.  Memory past n holds garbage, as the workspace does.  Every scan
shape the kernel is built for (SR 5..8), with one and with two passes, must
give the sequential forward and back substitution."""
import numpy as np
import pytest

W = 64

def dpp(v, old, ctrl, rmask=0xF):
    out = old.copy()
    for i in range(W):
        row = i // 16
        if not (rmask >> row) & 1:
            continue
        src = None
        if 0x111 <= ctrl <= 0x11F:      # row_shr n
            n = ctrl - 0x110
            if (i % 16) >= n: src = i - n
        elif ctrl == 0x142:            # row_bcast 15
            if row >= 1: src = (row - 1) * 16 + 15
        elif ctrl == 0x143:            # row_bcast 31
            if row >= 2: src = 31
        elif ctrl == 0x138:            # wave_shr 1
            if i >= 1: src = i - 1
        if src is not None:
            out[i] = v[src]
    return out

def affine_step(a, b, ctrl, rmask):
    ap = dpp(a, np.ones(W), ctrl, rmask); bp = dpp(b, np.zeros(W), ctrl, rmask)
    return a * ap, a * bp + b

def affine_step_shr(a, b, ctrl):
    ap = dpp(a, np.ones(W), ctrl); bp = dpp(b, np.zeros(W), ctrl)   # zero-filled b'
    return a * ap, a * bp + b

def affine_scan(a, b):
    for c in (0x111, 0x112, 0x114, 0x118): a, b = affine_step_shr(a, b, c)
    a, b = affine_step(a, b, 0x142, 0xA)
    a, b = affine_step(a, b, 0x143, 0xC)
    return a, b

def pass_scan(al, be, carry):
    SR = al.shape[1]
    al = al.copy(); be = be.copy()
    for i in range(1, SR):
        be[:, i] = al[:, i] * be[:, i - 1] + be[:, i]; al[:, i] = al[:, i] * al[:, i - 1]
    A, B = affine_scan(al[:, SR - 1].copy(), be[:, SR - 1].copy())
    last = A * carry + B
    tin = dpp(last, np.full(W, carry), 0x138)
    return al, be, tin, last[W - 1]

def wave_subst(l1, d, u1, l2c3, u2c1, b, n, SR, garbage):
    span = W * SR
    L = len(l1)
    bv = np.concatenate([b, garbage[: L - n]])   # memory past n holds garbage
    carry = 0.0
    for p0 in range(0, n, span):
        j0 = p0 + SR * np.arange(W)
        idx = j0[:, None] + np.arange(SR)[None, :]
        al = -l1[idx]; be = bv[idx].copy()
        corr = l2c3 * (al[:, 1] * be[:, 0] + be[:, 1])
        if p0 == 0: be[0, 3] -= corr[0]
        al, be, tin, carry = pass_scan(al, be, carry)
        x = al * tin[:, None] + be
        for lane in range(W):
            for i in range(SR):
                if idx[lane, i] < n: bv[idx[lane, i]] = x[lane, i]
    carry = 0.0
    for p in range((n - 1) // span, -1, -1):
        jt = p * span + span - 1 - SR * np.arange(W)
        idx = jt[:, None] - np.arange(SR)[None, :]
        inn = idx < n
        al = np.where(inn, -d[idx] * u1[idx], 0.0); be = np.where(inn, d[idx] * bv[idx], 0.0)
        y0, y1 = bv[0], bv[1]
        al2, be2, tin, carry = pass_scan(al, be, carry)
        x = al2 * tin[:, None] + be2
        low = p == 0
        if low:
            xl = x[W - 1]
            x1 = ((y1 - u2c1 * xl[SR - 4]) - u1[1] * xl[SR - 3]) * d[1]
            x0 = (y0 - u1[0] * x1) * d[0]
            x[W - 1, SR - 2] = x1; x[W - 1, SR - 1] = x0
        for lane in range(W):
            for i in range(SR):
                if idx[lane, i] < n: bv[idx[lane, i]] = x[lane, i]
    return bv[:n]

def reference(l1, d, u1, l2c3, u2c1, b, n):
    # forward: y_j = b_j - l1_j y_{j-1}, row 3 also - l2 y_1; back: x_j = d_j (y_j - u1_j x_{j+1}), row 1 also - u2 x_3
    y = np.zeros(n)
    for j in range(n):
        y[j] = b[j] - (l1[j] * y[j - 1] if j else 0.0) - (l2c3 * y[1] if j == 3 else 0.0)
    x = np.zeros(n)
    for j in range(n - 1, -1, -1):
        x[j] = d[j] * (y[j] - (u1[j] * x[j + 1] if j + 1 < n else 0.0) - (u2c1 * x[3] if j == 1 else 0.0))
    return x



@pytest.mark.parametrize("n", [129, 323, 473, 640])
@pytest.mark.parametrize("SR", [5, 6, 7, 8])
def test_wavefront_substitution_model(n, SR):
    rng = np.random.default_rng(1000 * n + SR)
    L = ((n + W * SR - 1) // (W * SR) + 1) * W * SR + 64
    l1 = rng.uniform(-1, 1, L)
    u1 = rng.uniform(-1, 1, L)
    d = rng.uniform(0.5, 2, L)
    u1[n - 1] = 0.0
    l1[0] = 0.0
    b = rng.uniform(-1, 1, n)
    garbage = rng.uniform(-1e6, 1e6, L)
    x = wave_subst(l1, d, u1, 0.3, -0.2, b, n, SR, garbage)
    r = reference(l1, d, u1, 0.3, -0.2, b, n)
    assert np.max(np.abs(x - r)) < 1e-14 * np.max(np.abs(r))
