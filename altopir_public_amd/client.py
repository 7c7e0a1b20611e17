"""Client of the scheme: secret and packing keys, queries, and decryption of responses.

Mirrors the reference's client helpers (src/service.rs:227-257):
  * new_request(params, rows, index) -> (Client, frame): a fresh client (secret key and
    packing keys, src/packing.rs:1935-1984 PackingKeys::init_full with generate_ksk_body,
    :2154-2175) and the request frame of a query for record `index` (src/client.rs:288-426
    generate_query_impl: one RLWE encryption per n records, the b rows concatenated), the query
    words switched from Z_Q to Z_2^32 as new_request does;
  * decrypt_response(params, client, bytes) -> record bytes: each output part recovered from
    its q2 / q1 fields (src/modulus_switch.rs:71-96 recover), decrypted with the secret
    (s * ct0 + ct1 mod Q) and rescaled to Z_256 (src/client.rs:109-127), raising ValueError on
    a response of the wrong length.
Ciphertexts follow the reference's convention: ct = (ct0, ct1) = (-a, a s + e + m), decryption
s ct0 + ct1.  The public polynomials a_i, the masks of the packing keys (the reference's W_SEED
and V_SEED) and the secrets come from this module's own seeded generators, and the NTT is the
engine's (slot j = evaluation at psi^(2 bitrev(j) + 1), psi the smallest primitive 2n-th root
of unity): the reference draws them from spiral-rs (a git dependency absent here), so bytes
equal to the reference client's are not the aim; what the responses of the engine decrypt to
is (tests/test_client.py, tests/test_offline_gpu.py).  The server's side of the same scheme --
the hint and the packing precomputation from these public values -- is offline.py.

Sizes: the query of a database of R records is R/n ciphertext rows (R a multiple of n); the
secret and the errors are centred discrete Gaussians of width 0.5 sqrt(2 pi) (standard
deviation 0.5, src/params.rs:97-101, "Xs = Xe = D(0.5)").
"""
from __future__ import annotations

from dataclasses import dataclass
from functools import lru_cache

import numpy as np

from . import workload

Q = workload.Q
P = 256                 # plaintext modulus (src/params.rs:95)
Q1, Q2 = 1 << 10, 1 << 18  # row-1 and row-0 switching moduli (src/params.rs:72-82)
PUBLIC_SEED = 0         # the public query polynomials (the reference's SEED_0)
W_SEED, V_SEED = 7, 8   # masks of the two packing keys (src/scheme.rs:24-31)
SIGMA = 0.5


@dataclass(frozen=True)
class Params:
    """The scheme's parameters for a database of `rows` records of `cols` bytes."""
    rows: int
    cols: int
    n: int = 2048
    tg: int = 4

    def __post_init__(self):
        assert self.n >= 16 and self.n & (self.n - 1) == 0 and (Q - 1) % (2 * self.n) == 0
        assert self.rows % self.n == 0 and self.cols % self.n == 0

    @property
    def kp(self) -> int:
        return self.tg * self.n

    @property
    def outputs(self) -> int:
        return self.cols // self.n

    @property
    def part_len(self) -> int:
        return workload.part_len_for(self.n)

    @property
    def response_len(self) -> int:
        return self.outputs * self.part_len

    @property
    def request_len(self) -> int:
        return workload.request_len(self.rows, self.kp)

    @property
    def delta(self) -> int:
        return Q // P

    @property
    def gadget_bits(self) -> int:
        return -(-32 // self.tg)


# ------------------------------------------------------------------ NTT (numpy) --

@lru_cache(maxsize=None)
def _tables(n: int):
    psi = workload.primitive_root_2n(n)
    logn = n.bit_length() - 1
    br = np.array([workload.bitrev(k, logn) for k in range(n)], dtype=np.int64)
    fwd = np.array([pow(psi, int(b), Q) for b in br], dtype=np.uint64)       # psi^bitrev(k)
    inv = workload.inverse_table(n, psi)                                       # psi^-bitrev(k) / 2
    return psi, br, fwd, inv


def ntt(a: np.ndarray) -> np.ndarray:
    """Forward negacyclic NTT over the last axis (values < Q), slot j = evaluation at
    psi^(2 bitrev(j) + 1): Cooley-Tukey stages with psi^bitrev(m + i)."""
    a = np.array(a, dtype=np.uint64, copy=True) % np.uint64(Q)
    n = a.shape[-1]
    _, _, fwd, _ = _tables(n)
    lead = a.shape[:-1]
    q = np.uint64(Q)
    m, t = 1, n
    while m < n:
        t //= 2
        v = a.reshape(*lead, m, 2, t)
        s = fwd[m:2 * m].reshape(m, 1)
        u = v[..., 0, :].copy()
        w = (v[..., 1, :] * s) % q
        v[..., 0, :] = (u + w) % q
        v[..., 1, :] = (u + q - w) % q
        m *= 2
    return a


def intt(a: np.ndarray) -> np.ndarray:
    """Inverse of ntt (the engine's Gentleman-Sande transform with the table psi^-bitrev(k)/2)."""
    a = np.array(a, dtype=np.uint64, copy=True) % np.uint64(Q)
    n = a.shape[-1]
    _, _, _, inv = _tables(n)
    lead = a.shape[:-1]
    q = np.uint64(Q)
    half = np.uint64((Q + 1) // 2)
    t = 1
    while t < n:
        h = n // (2 * t)
        v = a.reshape(*lead, h, 2, t)
        w = (inv[h:2 * h] & np.uint64(0xFFFFFFFF)).reshape(h, 1)
        x = v[..., 0, :].copy()
        y = v[..., 1, :].copy()
        v[..., 0, :] = ((x + y) % q) * half % q
        v[..., 1, :] = ((x + q - y) % q) * w % q
        t *= 2
    return a


def auto_slots(t: int, n: int) -> np.ndarray:
    """tau_t in the NTT domain: tau_t(a)[j] = a[table[j]] (src/packing.rs:1082-1126)."""
    _, br, _, _ = _tables(n)
    e = (t * (2 * br + 1)) % (2 * n)
    return br[(e - 1) // 2]


def gen_pows(n: int) -> np.ndarray:
    """5^i mod 2n, i < n (src/packing.rs:1695-1703 with num_to_pack = n)."""
    g = np.ones(n, dtype=np.int64)
    for i in range(1, n):
        g[i] = g[i - 1] * 5 % (2 * n)
    return g


def negacyclic_mul(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """a * b in Z_Q[x]/(x^n + 1) (coefficients), through the NTT."""
    return intt(ntt(a) * ntt(b) % np.uint64(Q))


# ------------------------------------------------------------- public values --

def public_polys(count: int, n: int, seed: int = PUBLIC_SEED) -> np.ndarray:
    """The public query polynomials a_i (coefficients, uniform in [0, Q)), count x n."""
    return np.random.default_rng([0x50554253, seed]).integers(0, Q, size=(count, n), dtype=np.uint64)


def masks(params: Params) -> tuple[np.ndarray, np.ndarray]:
    """The packing keys' public masks w and v (tg x n NTT slots, uniform in [0, Q))."""
    w = np.random.default_rng([0x4D41534B, W_SEED]).integers(0, Q, size=(params.tg, params.n), dtype=np.uint64)
    v = np.random.default_rng([0x4D41534B, V_SEED]).integers(0, Q, size=(params.tg, params.n), dtype=np.uint64)
    return w, v


def hint_query_polys(params: Params, seed: int = PUBLIC_SEED) -> np.ndarray:
    """The polynomials the server's hint multiplies the records with: negacyclic_perm of each
    query ciphertext's first row -a_i (src/server.rs:711-737), coefficients, R/n x n."""
    a = public_polys(params.rows // params.n, params.n, seed)
    neg = (np.uint64(Q) - a) % np.uint64(Q)            # -a_i
    out = np.empty_like(neg)
    out[:, 0] = neg[:, 0]
    out[:, 1:] = (np.uint64(Q) - neg[:, :0:-1]) % np.uint64(Q)   # sigma: x^j -> -x^(n-j)
    return out


# ----------------------------------------------------------------- sampling --

@lru_cache(maxsize=None)
def _gauss_cdf():
    xs = np.arange(-8, 9)
    p = np.exp(-(xs.astype(np.float64) ** 2) / (2 * SIGMA ** 2))
    return xs, np.cumsum(p / p.sum())


def gaussian(rng: np.random.Generator, shape) -> np.ndarray:
    """Centred discrete Gaussian of standard deviation 0.5 (exact table over |x| <= 8)."""
    xs, cdf = _gauss_cdf()
    return xs[np.searchsorted(cdf, rng.random(shape), side="right").clip(0, len(xs) - 1)]


def _mod(x: np.ndarray) -> np.ndarray:
    return (np.asarray(x, dtype=np.int64) % Q).astype(np.uint64)


# ------------------------------------------------------------------- client --

@dataclass
class Client:
    params: Params
    s: np.ndarray          # secret, small signed coefficients (int64, n)
    y: np.ndarray          # packing key for tau_5 (tg x n NTT slots)
    z: np.ndarray          # packing key for tau_{2n-1}

    @property
    def s_ntt(self) -> np.ndarray:
        return ntt(_mod(self.s))


def ksk_body(params: Params, s: np.ndarray, t: int, mask: np.ndarray, rng: np.random.Generator) -> np.ndarray:
    """Key-switching key from tau_t(s) to s (src/packing.rs:2154-2175):
    -s * mask + e + tau_t(s) * B^k, k < tg, B = 2^gadget_bits, in NTT slots."""
    n, tg = params.n, params.tg
    s_hat = ntt(_mod(s))
    tau_s = s_hat[auto_slots(t, n)]
    e_hat = ntt(_mod(gaussian(rng, (tg, n))))
    q = np.uint64(Q)
    out = np.empty((tg, n), dtype=np.uint64)
    for k in range(tg):
        g = np.uint64(pow(2, params.gadget_bits * k, Q))
        out[k] = ((q - mask[k] % q) * s_hat % q + e_hat[k] + tau_s * g % q) % q
    return out


def new_client(params: Params, seed: int | None = None) -> Client:
    """A fresh secret and its two packing keys (src/packing.rs:1935-1984): y switches
    tau_5(s) (the generator of the rotations, gen_pows[1]) and z tau_{2n-1}(s) to s."""
    rng = np.random.default_rng(seed)
    s = gaussian(rng, params.n)
    w, v = masks(params)
    y = ksk_body(params, s, 5, w, rng)
    z = ksk_body(params, s, 2 * params.n - 1, v, rng)
    return Client(params, s, y, z)


def query(client: Client, index: int, seed: int | None = None) -> np.ndarray:
    """Query words (u32) for record `index`: row i of the query is a_i s + e_i + [i == index / n]
    Delta x^(index mod n) (src/client.rs:288-426), switched to Z_2^32 by
    round(v 2^32 / Q) (src/service.rs:233-236)."""
    p = client.params
    assert 0 <= index < p.rows
    rng = np.random.default_rng(seed)
    npolys = p.rows // p.n
    a = public_polys(npolys, p.n)
    b = intt(ntt(a) * client.s_ntt % np.uint64(Q))
    b = (b + _mod(gaussian(rng, (npolys, p.n)))) % np.uint64(Q)
    i0, j0 = divmod(index, p.n)
    b[i0, j0] = (b[i0, j0] + np.uint64(p.delta)) % np.uint64(Q)
    v = b.reshape(-1)
    return (((v << np.uint64(32)) + np.uint64(Q // 2)) // np.uint64(Q) & np.uint64(0xFFFFFFFF)).astype(np.uint32)


def new_request(params: Params, index: int, seed: int | None = None) -> tuple[Client, np.ndarray]:
    """src/service.rs:227-239: a fresh client and the request frame for record `index`."""
    client = new_client(params, seed)
    q = query(client, index, None if seed is None else seed + 1)
    frame = workload.encode_request(client.y.reshape(-1).astype(np.uint32), client.z.reshape(-1).astype(np.uint32), q)
    return client, frame


def rescale(a: np.ndarray, inp: int, out: int) -> np.ndarray:
    """Centred rounding of a in Z_inp to Z_out, half away from zero (src/modulus_switch.rs:120-134)."""
    v = np.asarray(a, dtype=np.int64) % inp
    v = np.where(v >= inp // 2, v - inp, v)
    num = v * out
    r = np.where(num >= 0, (num + inp // 2) // inp, -((-num + inp // 2) // inp))
    return (r % out).astype(np.uint64)


def read_fields(buf: np.ndarray, bit0: int, bits: int, count: int) -> np.ndarray:
    """count LSB-first fields of `bits` bits from bit offset bit0 (src/bits.rs)."""
    allbits = np.unpackbits(np.asarray(buf, dtype=np.uint8), bitorder="little")
    f = allbits[bit0:bit0 + bits * count].reshape(count, bits).astype(np.uint64)
    return (f << np.arange(bits, dtype=np.uint64)).sum(axis=1).astype(np.uint64)


def decrypt_part(client: Client, part: np.ndarray) -> np.ndarray:
    """One output part: row 0 (q2 fields) and row 1 (q1 fields) lifted to Z_Q, s ct0 + ct1,
    rescaled to Z_256 (src/service.rs:241-257, src/client.rs:109-127)."""
    n = client.params.n
    ct0 = rescale(read_fields(part, 0, 18, n), Q2, Q)
    ct1 = rescale(read_fields(part, 18 * n, 10, n), Q1, Q)
    v = (intt(ntt(ct0) * client.s_ntt % np.uint64(Q)) + ct1) % np.uint64(Q)
    return rescale(v, Q, P)


def decrypt_response(params: Params, client: Client, data) -> bytes:
    """The record a response encrypts (src/service.rs:241-257)."""
    buf = np.frombuffer(bytes(data), dtype=np.uint8)
    if len(buf) != params.response_len:
        raise ValueError(f"response of {len(buf)} bytes, expected {params.response_len}")
    out = []
    for o in range(params.outputs):
        m = decrypt_part(client, buf[o * params.part_len:(o + 1) * params.part_len])
        if (m > 255).any():
            raise ValueError("decrypted value out of range")
        out.append(m.astype(np.uint8))
    return np.concatenate(out).tobytes()
