"""Seeded synthetic cohorts of biallelic hard calls.

Genotype codes follow the unphased VCF index of the reference
(coefficients.rs:18-38): 0 = 0/0, 1 = 0/1, 2 = 1/1, and 3 marks a missing
call (vcf.rs:355,397 turn missing data into NaN, which drops the site for
every pair that includes the sample, lanevector.rs:100-104).

Allele frequencies p_s (of the REF allele) are Beta(0.5, 0.5) clamped to
[maf, 1 - maf]; founders are drawn under Hardy-Weinberg, a fraction of them
inbred (both alleles one identical-by-descent draw with probability F per
site); relatives are made by gene-dropping with independent sites into
parent-offspring, full-sib, half-sib and first-cousin pairs.
"""
from __future__ import annotations

import dataclasses

import numpy as np

REL_TYPES = ("PO", "FS", "HS", "FC")


@dataclasses.dataclass
class Cohort:
    codes: np.ndarray  # (N, V) uint8, sample-major
    p: np.ndarray  # (V,) generating REF-allele frequency
    pairs: list  # [(x, y, relationship)] designated relative pairs
    inbred: np.ndarray  # (N,) bool, founder drawn with F > 0
    names: list

    @property
    def n_samples(self) -> int:
        return self.codes.shape[0]

    @property
    def n_sites(self) -> int:
        return self.codes.shape[1]


def _founder(rng, q, f_inbred):
    h0 = (rng.random(q.shape[0]) < q).astype(np.uint8)
    h1 = (rng.random(q.shape[0]) < q).astype(np.uint8)
    if f_inbred > 0:
        ibd = rng.random(q.shape[0]) < f_inbred
        h1 = np.where(ibd, h0, h1)
    return np.stack([h0, h1])


def _child(rng, mother, father):
    v = mother.shape[1]
    pick_m = rng.integers(0, 2, v, dtype=np.uint8)
    pick_f = rng.integers(0, 2, v, dtype=np.uint8)
    idx = np.arange(v)
    return np.stack([mother[pick_m, idx], father[pick_f, idx]])


def generate(n_samples: int, n_sites: int, seed: int, maf: float = 0.01, family_fraction: float = 0.2,
             inbred_fraction: float = 0.02, f_inbred: float = 0.125, missing: float = 0.01,
             missing_per_sample_max: float | None = None, n_family_pairs: int | None = None) -> Cohort:
    """Generate a cohort.

    n_family_pairs fixes the number of designated relative pairs (cycled over
    PO/FS/HS/FC); otherwise pairs are added until family_fraction of the
    samples belong to one.  missing_per_sample_max switches missingness from a
    fixed MCAR rate to a per-sample rate ~ Uniform(0, max).
    """
    rng = np.random.default_rng(seed)
    p = rng.beta(0.5, 0.5, n_sites)
    p = np.clip(p, maf, 1.0 - maf)
    q = 1.0 - p

    haps = []  # list of (2, V) haplotype arrays of output samples
    inbred = []
    pairs = []

    def founder():
        inb = rng.random() < inbred_fraction
        return _founder(rng, q, f_inbred if inb else 0.0), inb

    def add(h, inb):
        haps.append(h)
        inbred.append(inb)
        return len(haps) - 1

    target = n_family_pairs
    if target is None:
        # each designated pair puts two output samples in a family
        target = int(round(family_fraction * n_samples / 2))
    k = 0
    while len(pairs) < target and len(haps) + 2 <= n_samples:
        kind = REL_TYPES[k % 4]
        k += 1
        if kind == "PO":
            (a, ia), (b, _) = founder(), founder()
            c = _child(rng, a, b)
            x, y = add(a, ia), add(c, False)
        elif kind == "FS":
            (a, _), (b, _) = founder(), founder()
            x, y = add(_child(rng, a, b), False), add(_child(rng, a, b), False)
        elif kind == "HS":
            (a, _), (b, _), (e, _) = founder(), founder(), founder()
            x, y = add(_child(rng, a, b), False), add(_child(rng, a, e), False)
        else:  # FC: children of two full sibs
            (g1, _), (g2, _) = founder(), founder()
            s1, s2 = _child(rng, g1, g2), _child(rng, g1, g2)
            (m1, _), (m2, _) = founder(), founder()
            x, y = add(_child(rng, s1, m1), False), add(_child(rng, s2, m2), False)
        pairs.append((x, y, kind))
    while len(haps) < n_samples:
        h, inb = founder()
        add(h, inb)

    codes = np.empty((n_samples, n_sites), dtype=np.uint8)
    for i, h in enumerate(haps):
        codes[i] = h[0] + h[1]
    if missing_per_sample_max is not None:
        rates = rng.uniform(0.0, missing_per_sample_max, n_samples)
    else:
        rates = np.full(n_samples, missing)
    for i in range(n_samples):
        if rates[i] > 0:
            codes[i, rng.random(n_sites) < rates[i]] = 3
    names = [f"S{i:05d}" for i in range(n_samples)]
    return Cohort(codes=codes, p=p, pairs=pairs, inbred=np.array(inbred), names=names)


# Configurations of the benchmark workloads (sizes and generator settings).
CONFIGS = {
    "c1": dict(n_samples=64, n_sites=10_000, seed=0xC1, maf=0.01, missing=0.01, inbred_fraction=0.0,
               n_family_pairs=16),
    "c2": dict(n_samples=1000, n_sites=100_000, seed=0xC2, maf=0.01, missing=0.01),
    "c3": dict(n_samples=4000, n_sites=50_000, seed=0xC3, maf=0.01, missing_per_sample_max=0.3),
    "c4": dict(n_samples=2000, n_sites=100_000, seed=0xC4, maf=0.005, missing=0.01),
    "c5": dict(n_samples=10_000, n_sites=100_000, seed=0xC5, maf=0.01, missing=0.01),
}


def config(name: str, **overrides) -> Cohort:
    kw = dict(CONFIGS[name])
    kw.update(overrides)
    return generate(**kw)


def soft_likelihoods(codes: np.ndarray, seed: int, mean_depth: float = 4.0, error: float = 0.01) -> np.ndarray:
    """Linear-scale genotype likelihoods (V, N, 3) of 0/0, 0/1, 1/1, site-major
    like the reference's GenotypeLikelihoods: each call gets Poisson(mean_depth)
    reads, each read shows the ALT allele with probability error, 1/2 or
    1 - error under the true genotype, and the binomial likelihoods are scaled
    by their maximum (vcf.rs:368 normalises GL the same way), so one entry is
    exactly 1.  A call with no read, or a missing code, is a NaN row."""
    rng = np.random.default_rng(seed)
    n, v = codes.shape
    depth = rng.poisson(mean_depth, (n, v))
    alt_p = np.array([error, 0.5, 1.0 - error, 0.0])[codes]
    alt = rng.binomial(depth, alt_p)
    ref = depth - alt
    # log P(reads | G) up to a constant, G = 0/0, 0/1, 1/1
    pa = np.array([error, 0.5, 1.0 - error])
    ll = alt[..., None] * np.log(pa) + ref[..., None] * np.log1p(-pa)
    ll -= ll.max(axis=-1, keepdims=True)
    gl = np.exp(ll)
    gl[(depth == 0) | (codes == 3)] = np.nan
    return np.ascontiguousarray(gl.transpose(1, 0, 2))


def multi_likelihoods(n_samples: int, n_sites: tuple, seed: int, mean_depth: float = 5.0, error: float = 0.01,
                      absent: float = 0.3, missing: float = 0.02):
    """Soft likelihood buckets (gl2, gl3, gl4) for sites with 2, 3, 4 alleles
    (each (V_A, N, A(A+1)/2), VCF genotype order), site-major like the
    reference's GenotypeLikelihoods.  Genotypes follow Hardy-Weinberg from
    Dirichlet allele frequencies; with probability `absent` one allele of a
    3/4-allele site is never carried (its fitted frequency is 0, so the site
    moves to a smaller bucket).  Reads: Poisson(mean_depth) per call, each the
    carried allele (one of the two at random) with probability 1 - error, else
    another allele; likelihoods scaled by their maximum; no read or a
    `missing` draw gives a NaN row."""
    rng = np.random.default_rng(seed)
    out = []
    for A, v in zip((2, 3, 4), n_sites):
        G = A * (A + 1) // 2
        gi = [i for j in range(A) for i in range(j + 1)]
        gj = [j for j in range(A) for i in range(j + 1)]
        gl = np.empty((v, n_samples, G))
        for s in range(v):
            p = rng.dirichlet(np.full(A, 0.7))
            if A > 2 and rng.random() < absent:
                p[rng.integers(A)] = 0.0
                p /= p.sum()
            alleles = rng.choice(A, (n_samples, 2), p=p)
            depth = rng.poisson(mean_depth, n_samples)
            ll = np.zeros((n_samples, G))
            for r in range(n_samples):
                for _ in range(depth[r]):
                    carried = alleles[r, rng.integers(2)]
                    seen = carried if rng.random() >= error else rng.choice([a for a in range(A) if a != carried])
                    for g in range(G):
                        pi = (1 - error) if seen == gi[g] else error / (A - 1)
                        pj = (1 - error) if seen == gj[g] else error / (A - 1)
                        ll[r, g] += np.log(0.5 * pi + 0.5 * pj)
            ll -= ll.max(axis=1, keepdims=True)
            row = np.exp(ll)
            row[(depth == 0) | (rng.random(n_samples) < missing)] = np.nan
            gl[s] = row
        out.append(gl)
    return tuple(out)


def write_vcf_multi(path: str, buckets, field: str = "GL", names=None) -> None:
    """A VCF of likelihood buckets (gl2, gl3, gl4) as multi_likelihoods makes
    them: records interleaved 2-, 3-, 4-allele site by site (each bucket keeps
    its order), GL as log10 with 7 significant digits or PL as Phred integers,
    '.' values for a NaN row."""
    import gzip

    n = next(b.shape[1] for b in buckets if b is not None and b.size)
    names = names or [f"S{i:05d}" for i in range(n)]
    alts = {2: "C", 3: "C,G", 4: "C,G,T"}

    def cells(rows):
        with np.errstate(divide="ignore"):
            lg = np.log10(rows)
        out = []
        for r in lg:
            if np.isnan(r).any():
                out.append(",".join(["."] * len(r)))
            elif field == "GL":
                out.append(",".join(f"{x:.7g}" for x in r))
            else:
                out.append(",".join(str(int(min(round(-10 * x), 99999))) for x in r))
        return "\t".join(out)

    opener = gzip.open if path.endswith(".gz") else open
    with opener(path, "wt") as f:
        f.write("##fileformat=VCFv4.2\n")
        f.write(f'##FORMAT=<ID={field},Number=G,Type={"Float" if field == "GL" else "Integer"},Description="">\n')
        f.write("#CHROM\tPOS\tID\tREF\tALT\tQUAL\tFILTER\tINFO\tFORMAT\t" + "\t".join(names) + "\n")
        pos = 0
        longest = max(b.shape[0] for b in buckets if b is not None)
        for k in range(longest):
            for A, b in zip((2, 3, 4), buckets):
                if b is None or k >= b.shape[0]:
                    continue
                pos += 1
                f.write(f"1\t{pos}\t.\tA\t{alts[A]}\t.\tPASS\t.\t{field}\t{cells(b[k])}\n")


def multi_genotypes(n_samples: int, n_sites: tuple, seed: int, family_fraction: float = 0.4,
                    missing: float = 0.0):
    """Hard calls at sites with 2, 3 and 4 alleles: one (N, V_A) uint8 array of
    genotype indices per bucket (VCF order, j(j+1)/2 + i for alleles i <= j;
    255 = a missing call), sample-major.  Allele frequencies are
    Dirichlet(0.7) per site; the first samples are founders drawing both
    alleles from them, and a `family_fraction` of the samples are children of
    two earlier samples (one allele of each parent, per site), so the cohort
    holds parent-child, sibling and more distant pairs.  A call is missing
    with probability `missing`."""
    rng = np.random.default_rng(seed)
    n_child = int(round(family_fraction * n_samples))
    n_found = max(2, n_samples - n_child)
    parents = [(int(rng.integers(k)), int(rng.integers(k))) for k in range(n_found, n_samples)]
    out = []
    for A, v in zip((2, 3, 4), n_sites):
        al = np.empty((n_samples, v, 2), np.int64)
        for s in range(v):
            p = rng.dirichlet(np.full(A, 0.7))
            al[:n_found, s] = rng.choice(A, (n_found, 2), p=p)
        for k, (mo, fa) in zip(range(n_found, n_samples), parents):
            pick_m = rng.integers(2, size=v)
            pick_f = rng.integers(2, size=v)
            al[k, :, 0] = al[mo, np.arange(v), pick_m]
            al[k, :, 1] = al[fa, np.arange(v), pick_f]
        lo, hi = al.min(axis=2), al.max(axis=2)
        idx = (hi * (hi + 1) // 2 + lo).astype(np.uint8)
        if missing > 0:
            idx[rng.random(idx.shape) < missing] = 255
        out.append(idx)
    return tuple(out)


def multi_onehot(idx_buckets):
    """One-hot likelihood rows of multi_genotypes' hard calls: per bucket a
    (V_A, N, A(A+1)/2) array, site-major as vcf.rs builds them, a missing
    call's row all NaN."""
    out = []
    for A, ix in zip((2, 3, 4), idx_buckets):
        n = ix.shape[0]
        rows = np.zeros((ix.shape[1], n, A * (A + 1) // 2))
        called = ix.T != 255
        s_i, n_i = np.nonzero(called)
        rows[s_i, n_i, ix.T[called]] = 1.0
        rows[~called] = np.nan
        out.append(rows)
    return tuple(out)


def write_vcf_gt_multi(path: str, buckets, names=None, af=None) -> None:
    """A GT VCF of genotype-index buckets (idx2, idx3, idx4) as multi_genotypes
    makes them: records interleaved 2-, 3-, 4-allele site by site (each bucket
    keeps its order), calls 'i/j' (or '.|.' for 255); af = (af2, af3, af4)
    writes INFO AF with the alternate frequencies (V_A x A arrays, 7
    significant digits)."""
    import gzip

    n = next(b.shape[0] for b in buckets if b is not None)
    names = names or [f"S{i:05d}" for i in range(n)]
    alts = {2: "C", 3: "C,G", 4: "C,G,T"}
    calls = {}
    for A in (2, 3, 4):
        t = [f"{i}/{j}" for j in range(A) for i in range(j + 1)]
        calls[A] = np.array(t + ["."] * (256 - len(t)), dtype=object)
        calls[A][255] = ".|."
    opener = gzip.open if path.endswith(".gz") else open
    with opener(path, "wt") as f:
        f.write("##fileformat=VCFv4.2\n")
        f.write('##INFO=<ID=AF,Number=A,Type=Float,Description="Allele frequency">\n')
        f.write('##FORMAT=<ID=GT,Number=1,Type=String,Description="Genotype">\n')
        f.write("#CHROM\tPOS\tID\tREF\tALT\tQUAL\tFILTER\tINFO\tFORMAT\t" + "\t".join(names) + "\n")
        pos = 0
        longest = max(b.shape[1] for b in buckets if b is not None)
        for k in range(longest):
            for bi, (A, b) in enumerate(zip((2, 3, 4), buckets)):
                if b is None or k >= b.shape[1]:
                    continue
                pos += 1
                inf = "AF=" + ",".join(f"{x:.7g}" for x in af[bi][k, 1:]) if af is not None else "."
                f.write(f"1\t{pos}\t.\tA\t{alts[A]}\t.\tPASS\t{inf}\tGT\t" + "\t".join(calls[A][b[:, k]]) + "\n")


def pack2(codes: np.ndarray, row_sites: int | None = None) -> np.ndarray:
    """Pack codes (N, V) into 2 bits per call, 4 calls per byte, site s of a
    row in bits 2*(s%4) of byte s//4.  Rows are padded with code 3 (missing)
    to row_sites sites (default: V rounded up to a multiple of 256)."""
    n, v = codes.shape
    if row_sites is None:
        row_sites = (v + 255) // 256 * 256
    padded = np.full((n, row_sites), 3, dtype=np.uint8)
    padded[:, :v] = codes
    r = padded.reshape(n, row_sites // 4, 4).astype(np.uint8)
    return (r[..., 0] | (r[..., 1] << 2) | (r[..., 2] << 4) | (r[..., 3] << 6)).astype(np.uint8)


def unpack2(packed: np.ndarray, n_sites: int) -> np.ndarray:
    b = packed.astype(np.uint8)
    out = np.stack([b & 3, (b >> 2) & 3, (b >> 4) & 3, (b >> 6) & 3], axis=-1).reshape(b.shape[0], -1)
    return out[:, :n_sites].copy()


class _BgzfWriter:
    """BGZF (htslib's blocked gzip): gzip members of at most 65,280 input bytes,
    each with a 'BC' extra field holding the block size, then the empty EOF
    block."""

    BLOCK = 65280

    def __init__(self, path):
        self.f = open(path, "wb")
        self.buf = bytearray()

    @staticmethod
    def _block(data: bytes) -> bytes:
        import struct
        import zlib

        c = zlib.compressobj(6, zlib.DEFLATED, -15)
        comp = c.compress(data) + c.flush()
        head = struct.pack("<BBBBIBBHBBHH", 0x1F, 0x8B, 8, 4, 0, 0, 0xFF, 6, ord("B"), ord("C"), 2, len(comp) + 25)
        return head + comp + struct.pack("<II", zlib.crc32(data) & 0xFFFFFFFF, len(data))

    def write(self, text: str):
        self.buf += text.encode()
        while len(self.buf) >= self.BLOCK:
            self.f.write(self._block(bytes(self.buf[:self.BLOCK])))
            del self.buf[:self.BLOCK]

    def close(self):
        if self.buf:
            self.f.write(self._block(bytes(self.buf)))
        self.f.write(self._block(b""))
        self.f.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def write_vcf(cohort: Cohort, path: str, field: str = "GT", af_info: bool = False, names=None,
              gl: np.ndarray | None = None, bgzf: bool = False) -> None:
    """Write the cohort as a VCF (plain text, or gzip when path ends in .gz).
    GT: 0/0, 0/1, 1/1, ./. for a missing call.  GL / PL: one-hot likelihoods
    (log10 likelihood 0 for the call and -1000 otherwise; Phred 0 and 10000),
    '.,.,.' for a missing call; or, with gl (V, N, 3) linear-scale
    likelihoods, their log10 (GL, 7 significant digits) or Phred-scaled
    integers (PL).  af_info writes INFO AF = 1 - p.  bgzf: BGZF blocks (as
    bgzip writes them) instead of one gzip stream."""
    import gzip

    codes = cohort.codes
    n, v = codes.shape
    names = names or cohort.names
    gt = np.array(["0/0", "0/1", "1/1", "./."])
    if field == "GL":
        lik = np.array(["0,-1000,-1000", "-1000,0,-1000", "-1000,-1000,0", ".,.,."])
    else:
        lik = np.array(["0,10000,10000", "10000,0,10000", "10000,10000,0", ".,.,."])
    if bgzf:
        def opener(p, mode):
            return _BgzfWriter(p)
    elif path.endswith(".gz"):
        def opener(p, mode):
            return gzip.open(p, mode, compresslevel=1)
    else:
        opener = open
    bases = "ACGT"
    with opener(path, "wt") as f:
        f.write("##fileformat=VCFv4.2\n")
        f.write('##INFO=<ID=AF,Number=A,Type=Float,Description="Allele frequency">\n')
        f.write('##FORMAT=<ID=GT,Number=1,Type=String,Description="Genotype">\n')
        f.write('##FORMAT=<ID=GL,Number=G,Type=Float,Description="Genotype likelihoods">\n')
        f.write('##FORMAT=<ID=PL,Number=G,Type=Integer,Description="Phred-scaled genotype likelihoods">\n')
        f.write("#CHROM\tPOS\tID\tREF\tALT\tQUAL\tFILTER\tINFO\tFORMAT\t" + "\t".join(names) + "\n")
        table = gt if field == "GT" else lik
        for s in range(v):
            ref, alt = bases[s % 4], bases[(s + 1) % 4]
            inf = f"AF={1.0 - cohort.p[s]:.6g}" if af_info else "."
            if gl is not None and field != "GT":
                with np.errstate(divide="ignore"):
                    lg = np.log10(gl[s])
                if field == "GL":
                    cells = [".,.,." if np.isnan(r).any() else ",".join(f"{x:.7g}" for x in r) for r in lg]
                else:
                    cells = [".,.,." if np.isnan(r).any() else ",".join(str(int(min(round(-10 * x), 99999)))
                                                                         for x in r) for r in lg]
                calls = "\t".join(cells)
            else:
                calls = "\t".join(table[codes[:, s]])
            f.write(f"1\t{s + 1}\t.\t{ref}\t{alt}\t.\tPASS\t{inf}\t{field}\t{calls}\n")
