"""CPU ORACLE -- test infrastructure only (tests/ import it; the product never does).

Float64 numpy restatement of the learned double-excitation proposal the GPU
engine implements (the reference has no counterpart: it is an addition beside
the heat-bath generator of src/excit_gen_heat_bath_mol.F90, which supplies its
training samples).

Model.  d_in = d_out = 2*nbasis, two hidden layers of 256 with GELU (erf form):
    x_k = n_k,  x_{nb+k} = n_k * eps_k          (n_k occupation of spin orbital k+1)
    z1 = x W1 + b1, a1 = gelu(z1), z2 = a1 W2 + b2, a2 = gelu(z2), z3 = a2 W3 + b3
Parameters are one float32 vector: W1[d_in][256], b1, W2[256][256], b2,
W3[256][d_out], b3.  Inference on the GPU uses bf16 weights and bf16 layer
inputs with fp32 accumulation; ``forward(..., bf16=True)`` emulates the
roundings.  The per-determinant proposal weights are
    e_k = exp(clip(z3_k, -60, 60))        hole weights   (k < nb)
    f_k = exp(clip(z3_{nb+k}, -60, 60))   particle weights
stored as float32; every probability below is computed from the stored values.

Hole-particle affinity.  A learned, determinant-independent table
theta[h][p] (nbasis x nbasis, after b3 in the parameter vector, zero at
initialisation) gives A[h][p] = exp(clip(theta, -60, 60)), stored as float32.
The network component draws an ordered hole pair (h1, h2), the first particle
p1 paired with h1, then its partner p2 paired with h2:
    P(h1, h2) = e_h1/Zh * e_h2/(Zh - e_h1)
    P(p1 | h1) = f_p1 A[h1][p1] / Z1(h1),           Z1(h) = sum_{x in U} f_x A[h][x]
    P(p2 | h2, p1) = f_p2 A[h2][p2] / Z2(h2, C),     Z2(h, C) = sum_{x in C} f_x A[h][x]
with C = C(p1) the unoccupied orbitals sharing p2's required spin and irrep,
excluding p1.  With A = 1 this is the product proposal of per-orbital hole and
particle weights.  For the unordered double {i,j} -> {a,b} the four paths
(h1, p1) in {i, j} x {a, b} sum to

    q_mlp = P(i,j) [f_a A_ia/Z1(i) f_b A_jb/Z2(j,C_b) + f_b A_ib/Z1(i) f_a A_ja/Z2(j,C_a)]
          + P(j,i) [f_a A_ja/Z1(j) f_b A_ib/Z2(i,C_b) + f_b A_jb/Z1(j) f_a A_ia/Z2(i,C_a)]

and the mixture with fraction ``mix`` of the uniform fallback is
    q = (1 - mix) q_mlp + mix q_unif
    q_unif = 2/(nel(nel-1)) * 1/nvirt * [1/|C_b| + 1/|C_a|]
with C_x the unoccupied orbitals the partner of x may come from (x excluded).

Training loss for samples (D, ex, w): L = sum_s w_s (-log q_mlp(ex_s|D_s)) / sum_s w_s.
A sample flagged flat (its determinant gets no network row in production) is
scored with e = f = 1 (logits 0): only the affinity receives its gradient.
"""
import numpy as np

HIDDEN = 256
CLIP = 60.0


def sizes(nb):
    d_in = d_out = 2 * nb
    o = {}
    off = 0
    for name, n in (("W1", d_in * HIDDEN), ("b1", HIDDEN), ("W2", HIDDEN * HIDDEN), ("b2", HIDDEN),
                    ("W3", HIDDEN * d_out), ("b3", d_out), ("TH", nb * nb)):
        o[name] = (off, n)
        off += n
    return d_in, d_out, o, off


def unpack(params, nb):
    d_in, d_out, o, _ = sizes(nb)
    p = np.asarray(params, dtype=np.float64)
    g = lambda k: p[o[k][0]:o[k][0] + o[k][1]]  # noqa: E731
    return (g("W1").reshape(d_in, HIDDEN), g("b1"), g("W2").reshape(HIDDEN, HIDDEN), g("b2"),
            g("W3").reshape(HIDDEN, d_out), g("b3"))


def theta(params, nb):
    """The hole-particle affinity logits, (nb, nb) [hole - 1][particle - 1]."""
    _d_in, _d_out, o, _ = sizes(nb)
    p = np.asarray(params, dtype=np.float64)
    return p[o["TH"][0]:o["TH"][0] + o["TH"][1]].reshape(nb, nb)


def affinity(params, nb):
    """The stored affinity table A = exp(clip(theta)), float32."""
    return np.exp(np.clip(theta(params, nb), -CLIP, CLIP)).astype(np.float32)


def pack(W1, b1, W2, b2, W3, b3, TH=None):
    nb = len(b3) // 2
    TH = np.zeros(nb * nb) if TH is None else np.asarray(TH).ravel()
    return np.concatenate([W1.ravel(), b1, W2.ravel(), b2, W3.ravel(), b3, TH])


def init_params(nb, seed):
    """Deterministic small random parameters (for tests)."""
    d_in, d_out, _, n = sizes(nb)
    rng = np.random.default_rng(seed)
    W1 = rng.normal(0, 1.0 / np.sqrt(d_in), (d_in, HIDDEN))
    W2 = rng.normal(0, 1.0 / np.sqrt(HIDDEN), (HIDDEN, HIDDEN))
    W3 = rng.normal(0, 1.0 / np.sqrt(HIDDEN), (HIDDEN, d_out))
    b1 = rng.normal(0, 0.1, HIDDEN)
    b2 = rng.normal(0, 0.1, HIDDEN)
    b3 = rng.normal(0, 0.1, d_out)
    TH = rng.normal(0, 0.5, (nb, nb))
    return pack(W1, b1, W2, b2, W3, b3, TH).astype(np.float32)


def bf16(x):
    """Round float32 values to bfloat16 (round to nearest even), returned as float64."""
    a = np.asarray(x, dtype=np.float32).copy()
    u = a.view(np.uint32).astype(np.uint64)
    rounded = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    nan = np.isnan(a)
    out = rounded.astype(np.uint32).view(np.float32).astype(np.float64)
    out[nan] = np.nan
    return out


def occupations(dets, nb):
    """dets: (n, W) uint64 words -> (n, nb) 0/1 occupations of spin orbitals 1..nb."""
    dets = np.asarray(dets, dtype=np.uint64).reshape(len(dets), -1)
    k = np.arange(nb)
    words = dets[:, k // 64]
    return ((words >> (k % 64).astype(np.uint64)) & np.uint64(1)).astype(np.float64)


def features(occ, eps):
    return np.concatenate([occ, occ * np.asarray(eps, dtype=np.float64)[None, :]], axis=1)


def _erf(x):
    from scipy.special import erf
    return erf(x)


def gelu(z):
    return 0.5 * z * (1.0 + _erf(z / np.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + _erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def forward(params, occ, eps, bf16_inference=False):
    """Logits z3 (n, d_out); with bf16_inference the weights and every layer
    input are rounded to bf16 (accumulation in float64 here, fp32 on the GPU)."""
    nb = occ.shape[1]
    W1, b1, W2, b2, W3, b3 = unpack(params, nb)
    x = features(occ, eps)
    if bf16_inference:
        W1, W2, W3 = bf16(W1), bf16(W2), bf16(W3)
        b1, b2, b3 = (np.asarray(b, np.float32).astype(np.float64) for b in (b1, b2, b3))
        x = bf16(x)
    z1 = x @ W1 + b1
    a1 = gelu(z1)
    if bf16_inference:
        a1 = bf16(a1)
    z2 = a1 @ W2 + b2
    a2 = gelu(z2)
    if bf16_inference:
        a2 = bf16(a2)
    return a2 @ W3 + b3


def proposal_weights(z3):
    """The stored per-determinant weights: exp of the clipped logits, float32."""
    return np.exp(np.clip(z3, -CLIP, CLIP)).astype(np.float32)


# ---- spin / symmetry classes -------------------------------------------------
def ms(o):
    """o: 1-based spin orbital; odd orbitals are alpha (+1)."""
    return 1 if o % 2 == 1 else -1


def second_class(occset, orb_sym, nb, i, j, x):
    """Unoccupied orbitals the partner of x may be chosen from (spin and irrep
    conserving), excluding x itself."""
    msb = ms(i) + ms(j) - ms(x)
    if abs(msb) != 1:
        return []
    symb = orb_sym[i - 1] ^ orb_sym[j - 1] ^ orb_sym[x - 1]
    return [b for b in range(1, nb + 1)
            if b not in occset and b != x and ms(b) == msb and orb_sym[b - 1] == symb]


def _aff_or_ones(aff, nb):
    return np.ones((nb, nb)) if aff is None else np.asarray(aff, dtype=np.float64)


def _paths(e, f, A, occset, occ_list, unocc, orb_sym, nb, i, j, a, b):
    """The four sampling paths of the network component for {i,j} -> {a,b}:
    (h1, h2, p1, p2, class of p2, probability)."""
    Zh = sum(e[o - 1] for o in occ_list)
    P = {(i, j): e[i - 1] / Zh * e[j - 1] / (Zh - e[i - 1]), (j, i): e[j - 1] / Zh * e[i - 1] / (Zh - e[j - 1])}
    cls = {a: second_class(occset, orb_sym, nb, i, j, a), b: second_class(occset, orb_sym, nb, i, j, b)}
    out = []
    for h1, h2 in ((i, j), (j, i)):
        z1 = sum(f[x - 1] * A[h1 - 1, x - 1] for x in unocc)
        for p1, p2 in ((a, b), (b, a)):
            c = cls[p1]
            z2 = sum(f[x - 1] * A[h2 - 1, x - 1] for x in c)
            t = P[(h1, h2)] * (f[p1 - 1] * A[h1 - 1, p1 - 1] / z1) * (f[p2 - 1] * A[h2 - 1, p2 - 1] / z2)
            out.append((h1, h2, p1, p2, c, t))
    return out


def q_double(wrow, occ_list, orb_sym, nb, i, j, a, b, mix, aff=None):
    """Mixture probability of generating the unordered double {i,j}->{a,b}
    (given that a double was attempted); aff: the stored affinity table
    (None: all ones)."""
    occset = set(occ_list)
    e = np.asarray(wrow[:nb], dtype=np.float64)
    f = np.asarray(wrow[nb:2 * nb], dtype=np.float64)
    A = _aff_or_ones(aff, nb)
    nel = len(occ_list)
    unocc = [o for o in range(1, nb + 1) if o not in occset]
    q_mlp = sum(t for *_x, t in _paths(e, f, A, occset, occ_list, unocc, orb_sym, nb, i, j, a, b))
    cb = second_class(occset, orb_sym, nb, i, j, a)
    ca = second_class(occset, orb_sym, nb, i, j, b)
    q_unif = 2.0 / (nel * (nel - 1)) / len(unocc) * (1.0 / len(cb) + 1.0 / len(ca))
    return (1.0 - mix) * q_mlp + mix * q_unif, q_mlp, q_unif


def enumerate_doubles(wrow, occ_list, orb_sym, nb, mix, aff=None):
    """Every double the mixture can produce with its exact probability, by
    summing over the sampling paths (hole pair, first particle, second
    particle) of both components; plus the probability of a null attempt."""
    occset = set(occ_list)
    e = np.asarray(wrow[:nb], dtype=np.float64)
    f = np.asarray(wrow[nb:2 * nb], dtype=np.float64)
    A = _aff_or_ones(aff, nb)
    nel = len(occ_list)
    unocc = [o for o in range(1, nb + 1) if o not in occset]
    Zh = sum(e[o - 1] for o in occ_list)
    out = {}
    null = 0.0

    def add(key, p):
        out[key] = out.get(key, 0.0) + p

    # network component: ordered i, j; a paired with i; b (a's partner class) paired with j
    for i in occ_list:
        for j in occ_list:
            if i == j:
                continue
            pij = e[i - 1] / Zh * e[j - 1] / (Zh - e[i - 1])
            z1 = sum(f[x - 1] * A[i - 1, x - 1] for x in unocc)
            for a in unocc:
                pa = f[a - 1] * A[i - 1, a - 1] / z1
                cb = second_class(occset, orb_sym, nb, i, j, a)
                if not cb:
                    null += (1 - mix) * pij * pa
                    continue
                z2 = sum(f[x - 1] * A[j - 1, x - 1] for x in cb)
                for b in cb:
                    key = (min(i, j), max(i, j), min(a, b), max(a, b))
                    add(key, (1 - mix) * pij * pa * f[b - 1] * A[j - 1, b - 1] / z2)
    # uniform component: unordered pair, a uniform over U, b uniform over a's partner class
    npair = nel * (nel - 1) // 2
    for x in range(len(occ_list)):
        for y in range(x + 1, len(occ_list)):
            i, j = occ_list[x], occ_list[y]
            for a in unocc:
                cb = second_class(occset, orb_sym, nb, i, j, a)
                if not cb:
                    null += mix / npair / len(unocc)
                    continue
                for b in cb:
                    add((i, j, min(a, b), max(a, b)), mix / npair / len(unocc) / len(cb))
    return out, null


# ---- training loss ---------------------------------------------------------------
def _nll_and_grad(z3, th, occ_list, orb_sym, nb, ex):
    """-log q_mlp of one sample and its gradient with respect to the logits z3
    (2 nb) and to the affinity rows of the sample's holes i and j (nb each)."""
    i, j, a, b = ex
    occset = set(occ_list)
    u = np.clip(z3[:nb], -CLIP, CLIP)
    v = np.clip(z3[nb:2 * nb], -CLIP, CLIP)
    inside = np.abs(z3[:2 * nb]) < CLIP
    e = np.exp(u)
    f = np.exp(v)
    A = np.exp(np.clip(th, -CLIP, CLIP))
    O = np.zeros(nb, bool)
    O[[o - 1 for o in occ_list]] = True
    U = ~O
    unocc = [o for o in range(1, nb + 1) if o not in occset]
    Zh = e[O].sum()
    paths = _paths(e, f, A, occset, occ_list, unocc, orb_sym, nb, i, j, a, b)
    q = sum(t for *_x, t in paths)

    def g_hole(first, second):
        g = -e * O / Zh
        mask = O.copy()
        mask[first - 1] = False
        g = g - e * mask / (Zh - e[first - 1])
        g[first - 1] += 1.0
        g[second - 1] += 1.0
        return g

    gu = np.zeros(nb)
    grow = {i: np.zeros(nb), j: np.zeros(nb)}  # d log q / d theta[h][.] for h = i, j
    for h1, h2, p1, p2, c, t in paths:
        r = t / q
        gu += r * g_hole(h1, h2)
        w1 = f * A[h1 - 1] * U
        g1 = -w1 / w1.sum()
        g1[p1 - 1] += 1.0
        m2 = np.zeros(nb, bool)
        m2[[o - 1 for o in c]] = True
        w2 = f * A[h2 - 1] * m2
        g2 = -w2 / w2.sum()
        g2[p2 - 1] += 1.0
        grow[h1] += r * g1
        grow[h2] += r * g2
    gv = grow[i] + grow[j]
    grad = -np.concatenate([gu, gv]) * inside
    gth_i = -grow[i] * (np.abs(th[i - 1]) < CLIP)
    gth_j = -grow[j] * (np.abs(th[j - 1]) < CLIP)
    return -np.log(q), grad, gth_i, gth_j


def loss_and_grad(params, occ, eps, orb_sym, samples, weights, flat=None, habs=None):
    """Weighted mean NLL of the samples' doubles under q_mlp, and its gradient
    with respect to the flat parameter vector (float64 forward and backward);
    flat: per-sample flags (scored with zero logits, no network gradient).
    With habs (|H_ij| per sample) the loss is instead the spawning variance
    sum_s w_s |H_s| / q_mlp(s) / sum_s w_s: with w_s = |H_s| / p(s) for
    samples drawn with probability p, an estimate of sum |H|^2 / q; its
    gradient is the NLL gradient weighted by w_s |H_s| / q_mlp(s)."""
    nb = occ.shape[1]
    W1, b1, W2, b2, W3, b3 = unpack(params, nb)
    th = theta(params, nb)
    x = features(occ, eps)
    z1 = x @ W1 + b1
    a1 = gelu(z1)
    z2 = a1 @ W2 + b2
    a2 = gelu(z2)
    z3 = a2 @ W3 + b3
    w = np.asarray(weights, dtype=np.float64)
    wt = w / w.sum()
    dz3 = np.zeros_like(z3)
    dth = np.zeros((nb, nb))
    loss = 0.0
    for s in range(len(samples)):
        if w[s] == 0.0:
            continue
        ol = [k + 1 for k in range(nb) if occ[s, k] > 0.5]
        is_flat = flat is not None and flat[s]
        nll, g, gi, gj = _nll_and_grad(np.zeros_like(z3[s]) if is_flat else z3[s], th, ol, orb_sym, nb, samples[s])
        if habs is None:
            c = wt[s]
            loss += c * nll
        else:  # d(w h / q) = (w h / q) d(-log q)
            c = wt[s] * habs[s] * np.exp(nll)
            loss += c
        dz3[s] = 0.0 if is_flat else c * g
        i, j = samples[s][0], samples[s][1]
        dth[i - 1] += c * gi
        dth[j - 1] += c * gj
    dW3 = a2.T @ dz3
    db3 = dz3.sum(0)
    dz2 = (dz3 @ W3.T) * gelu_grad(z2)
    dW2 = a1.T @ dz2
    db2 = dz2.sum(0)
    dz1 = (dz2 @ W2.T) * gelu_grad(z1)
    dW1 = x.T @ dz1
    db1 = dz1.sum(0)
    return loss, pack(dW1, db1, dW2, db2, dW3, db3, dth)


def loss_only(params, occ, eps, orb_sym, samples, weights, flat=None, habs=None):
    nb = occ.shape[1]
    z3 = forward(params, occ, eps)
    th = theta(params, nb)
    w = np.asarray(weights, dtype=np.float64)
    wt = w / w.sum()
    total = 0.0
    for s in range(len(samples)):
        if w[s] == 0.0:
            continue
        ol = [k + 1 for k in range(nb) if occ[s, k] > 0.5]
        zs = np.zeros_like(z3[s]) if (flat is not None and flat[s]) else z3[s]
        nll = _nll_and_grad(zs, th, ol, orb_sym, nb, samples[s])[0]
        total += wt[s] * (nll if habs is None else habs[s] * np.exp(nll))
    return total
