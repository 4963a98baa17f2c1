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

Proposal for a double excitation {i,j} -> {a,b} from determinant D (occupied
set O, unoccupied U), mixture with fraction ``mix`` of the uniform fallback:
    q = (1 - mix) q_mlp + mix q_unif
    q_mlp = [e_i/Zh * e_j/(Zh-e_i) + e_j/Zh * e_i/(Zh-e_j)]
            * [f_a/Zp * f_b/Z(C_b \\ a) + f_b/Zp * f_a/Z(C_a \\ b)]
    q_unif = 2/(nel(nel-1)) * 1/nvirt * [1/|C_b \\ a| + 1/|C_a \\ b|]
with Zh = sum_O e, Zp = sum_U f, C_x the unoccupied orbitals sharing x's spin
and irrep (the class the second particle must come from), Z(.) sums of f.

Training loss for samples (D, ex, w): L = sum_s w_s (-log q_mlp(ex_s|D_s)) / sum_s w_s.
"""
import numpy as np

HIDDEN = 256
CLIP = 60.0


def sizes(nb):
    d_in = d_out = 2 * nb
    o = {}
    off = 0
    for name, n in (("W1", d_in * HIDDEN), ("b1", HIDDEN), ("W2", HIDDEN * HIDDEN), ("b2", HIDDEN),
                    ("W3", HIDDEN * d_out), ("b3", d_out)):
        o[name] = (off, n)
        off += n
    return d_in, d_out, o, off


def unpack(params, nb):
    d_in, d_out, o, _ = sizes(nb)
    p = np.asarray(params, dtype=np.float64)
    g = lambda k: p[o[k][0]:o[k][0] + o[k][1]]  # noqa: E731
    return (g("W1").reshape(d_in, HIDDEN), g("b1"), g("W2").reshape(HIDDEN, HIDDEN), g("b2"),
            g("W3").reshape(HIDDEN, d_out), g("b3"))


def pack(W1, b1, W2, b2, W3, b3):
    return np.concatenate([W1.ravel(), b1, W2.ravel(), b2, W3.ravel(), b3])


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
    return pack(W1, b1, W2, b2, W3, b3).astype(np.float32)


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


def q_double(wrow, occ_list, orb_sym, nb, i, j, a, b, mix):
    """Mixture probability of generating the unordered double {i,j}->{a,b}
    (given that a double was attempted)."""
    occset = set(occ_list)
    e = np.asarray(wrow[:nb], dtype=np.float64)
    f = np.asarray(wrow[nb:2 * nb], dtype=np.float64)
    nel = len(occ_list)
    unocc = [o for o in range(1, nb + 1) if o not in occset]
    Zh = sum(e[o - 1] for o in occ_list)
    Zp = sum(f[o - 1] for o in unocc)
    hp = e[i - 1] / Zh * e[j - 1] / (Zh - e[i - 1]) + e[j - 1] / Zh * e[i - 1] / (Zh - e[j - 1])
    cb = second_class(occset, orb_sym, nb, i, j, a)
    ca = second_class(occset, orb_sym, nb, i, j, b)
    zb = sum(f[o - 1] for o in cb)
    za = sum(f[o - 1] for o in ca)
    pp = f[a - 1] / Zp * f[b - 1] / zb + f[b - 1] / Zp * f[a - 1] / za
    q_mlp = hp * pp
    q_unif = 2.0 / (nel * (nel - 1)) / len(unocc) * (1.0 / len(cb) + 1.0 / len(ca))
    return (1.0 - mix) * q_mlp + mix * q_unif, q_mlp, q_unif


def enumerate_doubles(wrow, occ_list, orb_sym, nb, mix):
    """Every double the mixture can produce with its exact probability, by
    summing over the sampling paths (hole pair, first particle, second
    particle) of both components; plus the probability of a null attempt."""
    occset = set(occ_list)
    e = np.asarray(wrow[:nb], dtype=np.float64)
    f = np.asarray(wrow[nb:2 * nb], dtype=np.float64)
    nel = len(occ_list)
    unocc = [o for o in range(1, nb + 1) if o not in occset]
    Zh = sum(e[o - 1] for o in occ_list)
    Zp = sum(f[o - 1] for o in unocc)
    out = {}
    null = 0.0

    def add(key, p):
        out[key] = out.get(key, 0.0) + p

    # MLP component: ordered i, j, then a, then b from a's partner class
    for i in occ_list:
        for j in occ_list:
            if i == j:
                continue
            pij = e[i - 1] / Zh * e[j - 1] / (Zh - e[i - 1])
            for a in unocc:
                pa = f[a - 1] / Zp
                cb = second_class(occset, orb_sym, nb, i, j, a)
                if not cb:
                    null += (1 - mix) * pij * pa
                    continue
                zb = sum(f[o - 1] for o in cb)
                for b in cb:
                    key = (min(i, j), max(i, j), min(a, b), max(a, b))
                    add(key, (1 - mix) * pij * pa * f[b - 1] / zb)
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
def _nll_and_grad(z3, occ_list, orb_sym, nb, ex):
    """-log q_mlp and its gradient with respect to the logits of one sample."""
    i, j, a, b = ex
    occset = set(occ_list)
    u = np.clip(z3[:nb], -CLIP, CLIP)
    v = np.clip(z3[nb:2 * nb], -CLIP, CLIP)
    inside = np.abs(z3) < CLIP
    e = np.exp(u)
    f = np.exp(v)
    O = np.zeros(nb, bool)
    O[[o - 1 for o in occ_list]] = True
    U = ~O
    Zh = e[O].sum()
    Zp = f[U].sum()
    pij = e[i - 1] / Zh * e[j - 1] / (Zh - e[i - 1])
    pji = e[j - 1] / Zh * e[i - 1] / (Zh - e[j - 1])
    hp = pij + pji

    def g_hole(first, second):
        g = -e * O / Zh
        mask = O.copy()
        mask[first - 1] = False
        g = g - e * mask / (Zh - e[first - 1])
        g[first - 1] += 1.0
        g[second - 1] += 1.0
        return g

    gu = (pij * g_hole(i, j) + pji * g_hole(j, i)) / hp
    cb = second_class(occset, orb_sym, nb, i, j, a)
    ca = second_class(occset, orb_sym, nb, i, j, b)
    mb = np.zeros(nb, bool)
    mb[[o - 1 for o in cb]] = True
    ma = np.zeros(nb, bool)
    ma[[o - 1 for o in ca]] = True
    zb = f[mb].sum()
    za = f[ma].sum()
    pab = f[a - 1] / Zp * f[b - 1] / zb
    pba = f[b - 1] / Zp * f[a - 1] / za
    pp = pab + pba

    def g_part(first, second, m, z):
        g = -f * U / Zp - f * m / z
        g[first - 1] += 1.0
        g[second - 1] += 1.0
        return g

    gv = (pab * g_part(a, b, mb, zb) + pba * g_part(b, a, ma, za)) / pp
    grad = -np.concatenate([gu, gv]) * inside
    return -np.log(hp) - np.log(pp), grad


def loss_and_grad(params, occ, eps, orb_sym, samples, weights):
    """Weighted mean NLL of the samples' doubles under q_mlp, and its gradient
    with respect to the flat parameter vector (float64 forward and backward)."""
    nb = occ.shape[1]
    W1, b1, W2, b2, W3, b3 = unpack(params, nb)
    x = features(occ, eps)
    z1 = x @ W1 + b1
    a1 = gelu(z1)
    z2 = a1 @ W2 + b2
    a2 = gelu(z2)
    z3 = a2 @ W3 + b3
    w = np.asarray(weights, dtype=np.float64)
    wt = w / w.sum()
    dz3 = np.zeros_like(z3)
    loss = 0.0
    for s in range(len(samples)):
        if w[s] == 0.0:
            continue
        ol = [k + 1 for k in range(nb) if occ[s, k] > 0.5]
        nll, g = _nll_and_grad(z3[s], ol, orb_sym, nb, samples[s])
        loss += wt[s] * nll
        dz3[s] = wt[s] * g
    dW3 = a2.T @ dz3
    db3 = dz3.sum(0)
    dz2 = (dz3 @ W3.T) * gelu_grad(z2)
    dW2 = a1.T @ dz2
    db2 = dz2.sum(0)
    dz1 = (dz2 @ W2.T) * gelu_grad(z1)
    dW1 = x.T @ dz1
    db1 = dz1.sum(0)
    return loss, pack(dW1, db1, dW2, db2, dW3, db3)


def loss_only(params, occ, eps, orb_sym, samples, weights):
    nb = occ.shape[1]
    z3 = forward(params, occ, eps)
    w = np.asarray(weights, dtype=np.float64)
    wt = w / w.sum()
    total = 0.0
    for s in range(len(samples)):
        if w[s] == 0.0:
            continue
        ol = [k + 1 for k in range(nb) if occ[s, k] > 0.5]
        total += wt[s] * _nll_and_grad(z3[s], ol, orb_sym, nb, samples[s])[0]
    return total
