"""CPU restatement of the cost model: unit features, the perceptron's
forward and backward passes and one Adam step, in float64 numpy.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the HIP
kernels in tagmatch_amd/csrc/cost_model.hip.  The engine never uses it.

The reference has no cost model (its scheduler flushes batches in arrival
order, front_end.cc:350-376), so nothing here restates reference code: the
network, its gradient and Adam are the textbook definitions, and
tests/test_cost_model_oracle.py pins the gradient against central finite
differences and the features against brute-force definitions.
"""
import numpy as np

N_FEAT = 16


def unpack(p, n_in, h, n_out):
    """Views of the flat parameter array: W1[n_in][h], b1[h], W2[h][h],
    b2[h], W3[h][n_out], b3[n_out]."""
    p = np.asarray(p, np.float64)
    o = 0
    out = []
    for shape in ((n_in, h), (h,), (h, h), (h,), (h, n_out), (n_out,)):
        n = int(np.prod(shape))
        out.append(p[o:o + n].reshape(shape))
        o += n
    assert o == p.size, (o, p.size)
    return out


def forward(p, x, shape):
    """Returns y and the activations (z1, a1, z2, a2)."""
    W1, b1, W2, b2, W3, b3 = unpack(p, *shape)
    x = np.asarray(x, np.float64)
    z1 = x @ W1 + b1
    a1 = np.maximum(z1, 0.0)
    z2 = a1 @ W2 + b2
    a2 = np.maximum(z2, 0.0)
    y = a2 @ W3 + b3
    return y, (z1, a1, z2, a2)


def loss_grad(p, x, t, shape):
    """Mean squared error over all outputs and its gradient (flat, same
    layout as p)."""
    W1, b1, W2, b2, W3, b3 = unpack(p, *shape)
    x = np.asarray(x, np.float64)
    t = np.asarray(t, np.float64)
    y, (z1, a1, z2, a2) = forward(p, x, shape)
    cnt = y.size
    d = y - t
    loss = float(np.sum(d * d) / cnt)
    dy = 2.0 * d / cnt
    gW3 = a2.T @ dy
    gb3 = dy.sum(0)
    da2 = dy @ W3.T
    dz2 = da2 * (z2 > 0)
    gW2 = a1.T @ dz2
    gb2 = dz2.sum(0)
    da1 = dz2 @ W2.T
    dz1 = da1 * (z1 > 0)
    gW1 = x.T @ dz1
    gb1 = dz1.sum(0)
    g = np.concatenate([a.ravel() for a in (gW1, gb1, gW2, gb2, gW3, gb3)])
    return loss, g


def adam_step(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015, Algorithm 1); returns (p, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


def popcount6(w):
    w = np.asarray(w, np.uint32).reshape(-1, 6)
    return np.unpackbits(w.view(np.uint8), axis=1).sum(1).astype(np.int64)


def unit_features(mask, or_words, n_members, member_pop_sum, queries):
    """Feature row of one (partition, batch) unit: mask and or_words are the
    partition's 6-word AND and OR of its members; queries (B x 6) is the
    batch."""
    f = np.zeros(N_FEAT, np.float64)
    mask = np.asarray(mask, np.uint32).reshape(6)
    orw = np.asarray(or_words, np.uint32).reshape(6)
    q = np.asarray(queries, np.uint32).reshape(-1, 6)
    npart = float(n_members)
    f[0] = np.log2(1.0 + npart) / 20.0
    f[1] = popcount6(mask)[0] / 64.0
    f[11] = member_pop_sum / npart / 64.0 if npart > 0 else 0.0
    f[12] = popcount6(orw)[0] / 192.0
    B = q.shape[0]
    if B and npart > 0:
        c = popcount6(q).astype(np.float64)
        o = popcount6(q & orw).astype(np.float64)
        f[2] = np.log2(1.0 + B) / 12.0
        f[3] = c.mean() / 192.0
        f[4] = np.sqrt(max(np.mean(c * c) - c.mean() ** 2, 0.0)) / 192.0
        f[5] = c.min() / 192.0
        f[6] = c.max() / 192.0
        f[7] = np.mean(c < 80)
        f[8] = np.mean((c >= 80) & (c < 110))
        f[9] = np.mean((c >= 110) & (c < 140))
        f[10] = np.mean(c >= 140)
        f[13] = o.mean() / 192.0
        f[14] = o.max() / 192.0
        f[15] = np.log2(1.0 + B * npart) / 32.0
    return f


def partition_stats(words, part_offsets):
    """Per partition (in partitioned order): OR of members, member count,
    total member popcount."""
    words = np.asarray(words, np.uint32).reshape(-1, 6)
    P = len(part_offsets) - 1
    orw = np.zeros((P, 6), np.uint32)
    cnt = np.zeros(P, np.int64)
    pop = np.zeros(P, np.int64)
    pc = popcount6(words) if len(words) else np.zeros(0, np.int64)
    for p in range(P):
        a, b = int(part_offsets[p]), int(part_offsets[p + 1])
        if b > a:
            orw[p] = np.bitwise_or.reduce(words[a:b], axis=0)
            pop[p] = pc[a:b].sum()
        cnt[p] = b - a
    return orw, cnt, pop
