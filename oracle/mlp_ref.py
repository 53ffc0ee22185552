"""CPU oracle of the learned move-ordering network -- TEST INFRASTRUCTURE ONLY.

numpy restatement of the network the GPU trains (908 -> 256 -> 32 -> 1,
ReLU, pairwise ranking loss softplus(s_neg - s_pos + o / PLAIN_PER_LOGIT),
o the plain ordering's score difference the pair carries in the upper 24
bits of its depth word), in two forms:
  * forward_backward(..., bf16=True): FP32 arithmetic with BF16 rounding at
    the points where the GPU feeds BF16 MFMA operands (weights W1, W2,
    activations h1, and the gradients dz2, dz1): the values the kernels'
    results are compared with;
  * bf16=False in float64: the plain mathematics, pinned by finite
    differences in the tests.
The network has no counterpart in the reference (whose move ordering is fixed
scores, src/zeta.cl:2059-2104); its features follow the reference's piece
encoding (src/types.h:94-100) and move packing (src/types.h:67-74).
"""
import numpy as np

NIN, NH1, NH2 = 908, 256, 32
NPARAMS = NIN * NH1 + NH1 + NH1 * NH2 + NH2 + NH2 + 1  # 240961
PAIR_DTYPE = np.dtype([("b", "<u8", (4,)), ("stm", "<u4"), ("mpos", "<u4"), ("mneg", "<u4"), ("depth", "<u4")])
PLAIN_PER_LOGIT = 100


def pair_offsets(pairs):
    """Plain score difference (negative minus positive) of each pair."""
    return (np.asarray(pairs["depth"], dtype=np.uint32).view(np.int32) >> 8).astype(np.float64)


def depth_word(depth, offset):
    return np.uint32(((int(offset) << 8) | (int(depth) & 0xFF)) & 0xFFFFFFFF)


def bf16(x):
    """Round FP32 to BF16 (nearest even) and back, elementwise."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).reshape(a.shape)


def unpack(params):
    p = np.asarray(params)
    o = 0
    W1 = p[o:o + NIN * NH1].reshape(NIN, NH1); o += NIN * NH1
    b1 = p[o:o + NH1]; o += NH1
    W2 = p[o:o + NH1 * NH2].reshape(NH1, NH2); o += NH1 * NH2
    b2 = p[o:o + NH2]; o += NH2
    W3 = p[o:o + NH2]; o += NH2
    b3 = p[o:o + 1]
    return W1, b1, W2, b2, W3, b3


def pack(W1, b1, W2, b2, W3, b3):
    return np.concatenate([W1.ravel(), b1, W2.ravel(), b2, W3, b3])


def features(b, stm, move):
    """Active input indices of (position planes b[0..3], side to move, move)."""
    b0, b1, b2, b3 = (int(x) for x in b)
    occ = b1 | b2 | b3
    rel = (lambda s: s ^ 56) if stm else (lambda s: s)
    idx = []
    while occ:
        sq = (occ & -occ).bit_length() - 1
        occ &= occ - 1
        col = (b0 >> sq) & 1
        typ = ((b1 >> sq) & 1) | (((b2 >> sq) & 1) << 1) | (((b3 >> sq) & 1) << 2)
        own = 0 if col == stm else 1
        idx.append((own * 6 + typ - 1) * 64 + rel(sq))
    idx.append(768 + rel(move & 63))
    idx.append(832 + rel((move >> 6) & 63))
    idx.append(896 + ((move >> 19) & 7) - 1)
    ct = (move >> 27) & 7
    if ct:
        idx.append(902 + ct - 1)
    return idx


def design_matrix(pairs):
    """X [2P x 908]: rows [0,P) cutoff moves, rows [P,2P) failed moves."""
    P = len(pairs)
    X = np.zeros((2 * P, NIN), dtype=np.float64)
    for i, r in enumerate(pairs):
        for k, m in enumerate((int(r["mpos"]), int(r["mneg"]))):
            X[i + k * P, features(r["b"], int(r["stm"]) & 1, m)] = 1.0
    return X


def forward_backward(params, pairs, bf16_points=True):
    """Mean pairwise loss and its gradient (flat, ABI layout)."""
    W1, b1, W2, b2, W3, b3 = unpack(params)
    X = design_matrix(pairs)
    P = len(pairs)
    if bf16_points:
        f = np.float32
        X = X.astype(f)
        W1b, W2b = bf16(W1), bf16(W2)
        z1 = X @ W1b + b1.astype(f)
        h1 = np.maximum(z1, 0).astype(f)
        h1b = bf16(h1)
        z2 = h1b @ W2b + b2.astype(f)
        h2 = np.maximum(z2, 0).astype(f)
        s = h2 @ W3.astype(f) + f(b3[0])
        d = s[P:] - s[:P] + (pair_offsets(pairs) / PLAIN_PER_LOGIT).astype(f)
        sig = 1 / (1 + np.exp(-d))
        loss = float(np.mean(np.logaddexp(0, d)))
        ds = np.concatenate([-sig, sig]).astype(f) / f(P)
        dz2 = (ds[:, None] * W3[None, :].astype(f)) * (h2 > 0)
        dz2b = bf16(dz2)
        gW3 = (ds[:, None] * h2).sum(0)
        gb3 = np.array([ds.sum()])
        gb2 = dz2.sum(0)
        gW2 = h1b.T @ dz2b
        dh1 = dz2b @ W2b.T
        dz1 = dh1 * (h1 > 0)
        gb1 = dz1.sum(0)
        gW1 = X.T @ bf16(dz1)
    else:
        z1 = X @ W1 + b1
        h1 = np.maximum(z1, 0)
        z2 = h1 @ W2 + b2
        h2 = np.maximum(z2, 0)
        s = h2 @ W3 + b3[0]
        d = s[P:] - s[:P] + pair_offsets(pairs) / PLAIN_PER_LOGIT
        sig = 1 / (1 + np.exp(-d))
        loss = float(np.mean(np.logaddexp(0, d)))
        ds = np.concatenate([-sig, sig]) / P
        dz2 = (ds[:, None] * W3[None, :]) * (h2 > 0)
        gW3 = (ds[:, None] * h2).sum(0)
        gb3 = np.array([ds.sum()])
        gb2 = dz2.sum(0)
        gW2 = h1.T @ dz2
        dz1 = (dz2 @ W2.T) * (h1 > 0)
        gb1 = dz1.sum(0)
        gW1 = X.T @ dz1
    return loss, pack(gW1, gb1, gW2, gb2, gW3, gb3)


def scores(params, pos_planes, stm, moves, bf16_points=True):
    """Network scores of moves from one position."""
    W1, b1, W2, b2, W3, b3 = unpack(params)
    X = np.zeros((len(moves), NIN), dtype=np.float32)
    for i, m in enumerate(moves):
        X[i, features(pos_planes, stm, int(m))] = 1.0
    if bf16_points:
        h1 = np.maximum(X @ bf16(W1) + b1, 0)
        h2 = np.maximum(bf16(h1) @ bf16(W2) + b2, 0)
    else:
        h1 = np.maximum(X @ W1 + b1, 0)
        h2 = np.maximum(h1 @ W2 + b2, 0)
    return h2 @ W3 + b3[0]
