"""Seeded synthetic workloads for the alignment hot path.

Sequences use the reference's int8 codes (A0 C1 N2 G3 T4,
impl/SequenceCharToUInt8.cpp:10-30).  Every generator takes an explicit seed
so that tests, the benchmark and the CPU baseline see identical inputs.

Workloads (sizes as in the benchmark definitions):
  * ``one_pair``       10 kb x 10 kb, 20 planted fragments (30-200 bp,
                       15% substitutions, 2% 1-5 bp indels).
  * ``batch``          4,096 pairs, lengths log-uniform on 1-50 kb, GC
                       U[0.40, 0.60], 0-30 fragments of 20-300 bp at 10-25%
                       substitutions + 1-3% indels.
  * ``long_masked``    500 kb x 500 kb at GC 0.45 with 30% lowercase-n runs,
                       a 3 kb repeat family (200 copies, 10% divergence) and
                       300 planted fragments.
  * ``random_pairs``   independent random fragment pairs (lengths log-uniform
                       on 38-2,000 bp, GC U[0.30, 0.70]) for the lambda/K
                       score distribution.
"""
import os

import numpy as np

CODE = {"A": 0, "C": 1, "N": 2, "G": 3, "T": 4}
LETTERS = np.frombuffer(b"ACNGT", dtype=np.uint8)
BASES = np.array([0, 1, 3, 4], dtype=np.int8)
SEED_BASE = 20261203


def random_seq(rng, n, gc=0.5):
    """i.i.d. bases with P(C) = P(G) = gc/2 and P(A) = P(T) = (1-gc)/2."""
    u = rng.random(n)
    a, c, g = (1 - gc) / 2, gc / 2, gc / 2
    idx = (u >= a).astype(np.int64) + (u >= a + c) + (u >= a + c + g)
    return BASES[idx]


def mutate(rng, frag, p_sub, p_indel, gc=0.5):
    """Substitutions (to a different base, probability p_sub per base) and
    indel events (probability p_indel per base, half insertions of 1-5 random
    bases before the base, half deletions of 1-5 bases from it)."""
    frag = np.asarray(frag, dtype=np.int8)
    n = len(frag)
    if n == 0:
        return frag.copy()
    ev = np.flatnonzero(rng.random(n) < p_indel)
    is_del = rng.random(len(ev)) < 0.5
    lens = rng.integers(1, 6, len(ev))
    keep = np.ones(n, dtype=bool)
    for i, ln in zip(ev[is_del], lens[is_del]):
        keep[i:i + ln] = False
    out = frag.copy()
    cls = np.array([0, 1, 0, 2, 3])[np.clip(out, 0, 4)]
    sub = (rng.random(n) < p_sub) & (out != 2)
    shifted = BASES[(cls + rng.integers(1, 4, n)) % 4]
    out = np.where(sub, shifted, out).astype(np.int8)
    ins_at = ev[~is_del]
    ins_len = lens[~is_del]
    if len(ins_at):
        pos = np.repeat(ins_at, ins_len)
        vals = random_seq(rng, int(ins_len.sum()), gc)
        out = np.insert(out, pos, vals)
        keep = np.insert(keep, pos, True)
    return out[keep].astype(np.int8)


def plant(rng, s1, s2, nfrag, lo, hi, p_sub, p_indel, gc=0.5):
    """Writes nfrag shared fragments into s1 and (mutated) into s2, in place."""
    for _ in range(nfrag):
        ln = int(rng.integers(lo, hi + 1))
        if ln >= len(s1) or ln >= len(s2):
            continue
        frag = random_seq(rng, ln, gc)
        p1 = int(rng.integers(0, len(s1) - ln + 1))
        s1[p1:p1 + ln] = frag
        m = mutate(rng, frag, p_sub, p_indel, gc)[: len(s2)]
        p2 = int(rng.integers(0, len(s2) - len(m) + 1))
        s2[p2:p2 + len(m)] = m


def one_pair(seed=SEED_BASE + 0, L1=10000, L2=10000, nfrag=20):
    rng = np.random.default_rng(seed)
    s1 = random_seq(rng, L1)
    s2 = random_seq(rng, L2)
    plant(rng, s1, s2, nfrag, 30, 200, 0.15, 0.02)
    return s1, s2


def log_uniform_int(rng, lo, hi, size=None):
    return np.exp(rng.uniform(np.log(lo), np.log(hi), size=size)).astype(np.int64)


def batch(seed=SEED_BASE + 1, npairs=4096, lo=1000, hi=50000, max_frag=30, meta=False):
    """List of (s1, s2) pairs (with meta: and the number of fragments
    planted in each pair)."""
    rng = np.random.default_rng(seed)
    pairs, planted = [], []
    for _ in range(npairs):
        L1 = int(log_uniform_int(rng, lo, hi))
        L2 = int(log_uniform_int(rng, lo, hi))
        gc = float(rng.uniform(0.40, 0.60))
        s1 = random_seq(rng, L1, gc)
        s2 = random_seq(rng, L2, gc)
        nfrag = int(rng.integers(0, max_frag + 1))
        plant(rng, s1, s2, nfrag, 20, 300, float(rng.uniform(0.10, 0.25)), float(rng.uniform(0.01, 0.03)), gc)
        pairs.append((s1, s2))
        planted.append(nfrag)
    return (pairs, planted) if meta else pairs


def long_masked(seed=SEED_BASE + 2, L=500000, gc=0.45, masked_frac=0.30, te_len=3000, te_copies=200,
                nfrag=300):
    """Returns (text1, text2): ASCII sequences with lowercase 'n' runs, as the
    masked genome regions that cut1Gap strips (controlLayer4.cpp:1385-1390)."""
    rng = np.random.default_rng(seed)
    te = random_seq(rng, te_len, gc)
    out = []
    for _ in range(2):
        s = random_seq(rng, L, gc)
        for _ in range(te_copies):
            c = mutate(rng, te, 0.10, 0.0, gc)[:te_len]
            p = int(rng.integers(0, L - len(c)))
            s[p:p + len(c)] = c
        out.append(s)
    s1, s2 = out
    # planted fragments separated by 25-2000 bp gaps
    p1 = int(rng.integers(0, 1000))
    p2 = int(rng.integers(0, 1000))
    for _ in range(nfrag):
        ln = int(rng.integers(20, 301))
        if p1 + ln >= L or p2 + ln >= L:
            break
        frag = random_seq(rng, ln, gc)
        s1[p1:p1 + ln] = frag
        m = mutate(rng, frag, 0.15, 0.0, gc)
        s2[p2:p2 + len(m)] = m[: max(0, L - p2)]
        p1 += ln + int(rng.integers(25, 2001))
        p2 += len(m) + int(rng.integers(25, 2001))
    texts = []
    for s in (s1, s2):
        t = LETTERS[s.astype(np.int64)].copy()
        covered = 0
        target = int(masked_frac * L)
        while covered < target:
            ln = int(log_uniform_int(rng, 20, 5000))
            p = int(rng.integers(0, L - ln))
            t[p:p + ln] = ord("n")
            covered += ln
        texts.append(t.tobytes().decode())
    return texts[0], texts[1]


def random_pairs(seed=SEED_BASE + 3, n=1000000, lo=38, hi=2000, gc_lo=0.30, gc_hi=0.70):
    """Concatenated random fragment pairs: (seqs, ref_off, ref_len, qry_off, qry_len)."""
    rng = np.random.default_rng(seed)
    l1 = log_uniform_int(rng, lo, hi, n).astype(np.int32)
    l2 = log_uniform_int(rng, lo, hi, n).astype(np.int32)
    gc = rng.uniform(gc_lo, gc_hi, n)
    tot = int(l1.sum() + l2.sum())
    # per-pair GC: draw uniforms once and threshold against the pair's composition
    u = rng.random(tot)
    gcs = np.repeat(np.stack([gc, gc], 1).ravel(), np.stack([l1, l2], 1).ravel())
    seqs = np.where(u < gcs / 2, 1, np.where(u < gcs, 3, np.where(u < gcs + (1 - gcs) / 2, 0, 4))).astype(np.int8)
    lens = np.stack([l1, l2], 1).ravel().astype(np.int64)
    offs = np.concatenate([[0], np.cumsum(lens)[:-1]])
    return seqs, offs[0::2].copy(), l1, offs[1::2].copy(), l2


def scoring_settings(seed=SEED_BASE + 3, n=6):
    """n (match, mismatch, open, extend) settings from the sweep ranges."""
    rng = np.random.default_rng(seed + 1000)
    out = []
    for _ in range(n):
        out.append((int(rng.integers(1, 4)), -int(rng.integers(1, 5)), -int(rng.integers(2, 7)),
                    -int(rng.integers(1, 4))))
    return out


def to_text(codes):
    return LETTERS[np.asarray(codes, dtype=np.int64)].tobytes().decode()


def encode(text):
    """seq_to_int8 (SequenceCharToUInt8.cpp:46-52): acgtu/ACGTU -> codes, else N."""
    lut = np.full(256, 2, dtype=np.int8)
    for ch, v in (("A", 0), ("C", 1), ("G", 3), ("T", 4), ("U", 4)):
        lut[ord(ch)] = v
        lut[ord(ch.lower())] = v
    return lut[np.frombuffer(text.encode(), dtype=np.uint8)]


# The score profiles of the reference's scoreMatrix folder (scoreMatrix/0,
# /1, /10, /70, /85, /100; one per GFF class: 0 unused, 1 transposable
# element, 10 intergenic, 70 gene body, 85 exon / UTR, 100 CDS):
# match, mismatch, open1, ext1, open2, ext2, zdrop.
WEIGHTED_PROFILES = {
    0: (0, 0, 0, 0, 0, 0, 1073741824),
    1: (2, -3, -4, -2, -45, 0, 10000),
    10: (20, -30, -40, -20, -450, 0, 5000),
    70: (140, -210, -280, -140, -3150, 0, 5000),
    85: (170, -255, -340, -170, -3825, 0, 5000),
    100: (200, -300, -400, -200, -4500, 0, 5000),
}


def gene_categories(rng, n):
    """GFF-like categories for n positions: intergenic (10) with gene bodies
    (70) holding exons (85) and CDS (100), and transposable elements (1)."""
    cat = np.full(n, 10, np.int16)
    pos = int(rng.integers(0, 1500))
    while pos < n:
        if rng.random() < 0.2:  # a transposable element
            ln = int(rng.integers(300, 3000))
            cat[pos:pos + ln] = 1
            pos += ln + int(rng.integers(200, 2000))
            continue
        glen = int(rng.integers(1000, 6000))
        cat[pos:pos + glen] = 70
        e = pos + int(rng.integers(0, 200))
        while e < pos + glen:
            el = int(rng.integers(60, 400))
            cat[e:min(e + el, pos + glen)] = 85
            c0 = e + int(rng.integers(0, 30))
            cat[c0:min(e + el - 10, pos + glen)] = 100
            e += el + int(rng.integers(80, 800))
        pos += glen + int(rng.integers(500, 4000))
    return cat


# -- maskGenome inputs ---------------------------------------------------------

def _wrap(s, width):
    return "\n".join(s[i:i + width] for i in range(0, len(s), width)) if s else ""


def repeat_genome(rng, n, families=40, fam_len=(200, 3000), copies=(3, 60), p_sub=0.02, gc=0.45):
    """One chromosome of n bases (upper-case ACGT bytes) made of random
    sequence with copies of `families` repeat elements planted at random
    places, each copy with substitutions -- the k-mer spectrum a jellyfish
    count of a repetitive plant genome shows (many unique k-mers, a tail of
    high-count ones)."""
    g = LETTERS[[0, 1, 3, 4]][rng.choice(4, n, p=[(1 - gc) / 2, gc / 2, gc / 2, (1 - gc) / 2])].copy()
    for _ in range(families):
        L = int(rng.integers(*fam_len))
        fam = LETTERS[[0, 1, 3, 4]][rng.integers(0, 4, L)]
        for _ in range(int(rng.integers(*copies))):
            at = int(rng.integers(0, max(1, n - L)))
            c = fam.copy()
            sub = rng.random(L) < p_sub
            c[sub] = LETTERS[[0, 1, 3, 4]][rng.integers(0, 4, int(sub.sum()))]
            g[at:at + L] = c[:n - at]
    return g


def kmer_counts(genome, k, min_count):
    """The distinct k-mers of genome (uint8 ACGT bytes) seen more than
    min_count times, as (kmers [m x k] uint8, counts), counted on the forward
    strand only, as the masking scan compares windows."""
    n = len(genome) - k + 1
    if n <= 0:
        return np.zeros((0, k), np.uint8), np.zeros(0, np.int64)
    code = np.zeros(n, np.uint64)
    b = ((genome.astype(np.uint64) >> np.uint64(1)) & np.uint64(3))
    for j in range(k):
        code = (code << np.uint64(2)) | b[j:j + n]
    u, cnt = np.unique(code, return_counts=True)
    keep = cnt > min_count
    u, cnt = u[keep], cnt[keep]
    out = np.zeros((len(u), k), np.uint8)
    letters = np.frombuffer(b"ACTG", np.uint8)
    for j in range(k):
        out[:, k - 1 - j] = letters[((u >> np.uint64(2 * j)) & np.uint64(3)).astype(np.int64)]
    return out, cnt


def dump_text(kmers, counts):
    """A jellyfish ``dump`` (fasta: ">count" then the k-mer)."""
    lines = []
    for km, c in zip(kmers, counts):
        lines.append(f">{int(c)}")
        lines.append(km.tobytes().decode())
    return "\n".join(lines) + "\n"


def mask_case(rng, small=True):
    """Random maskGenome inputs as file texts: genome, dump, gff, sam, cds
    fasta, plus command options, exercising the reader and writer rules the
    reference has (mixed case, N runs and other bytes, header forms, k-mer
    records of other lengths, CDS ranges past a sequence end, SAM records
    below the identity threshold)."""
    nchr = int(rng.integers(1, 5))
    k = int(rng.choice([1, 2, 3, 5, 8, 12, 20, 31, 32, 37] if small else [20]))
    chroms = []
    for c in range(nchr):
        L = int(rng.integers(max(k, 1), 2500 if small else 200000))
        s = np.frombuffer(b"ACGT", np.uint8)[rng.integers(0, 4, L)].copy()
        if L > 60 and rng.random() < 0.7:  # a repeat inside the chromosome
            a, m = int(rng.integers(0, L // 2)), int(rng.integers(5, L // 4))
            b = int(rng.integers(L // 2, L - m))
            s[b:b + m] = s[a:a + m]
        for _ in range(int(rng.integers(0, 4))):  # N runs, lower case runs, other bytes
            a = int(rng.integers(0, L))
            m = int(rng.integers(1, 40))
            kind = rng.random()
            if kind < 0.4:
                s[a:a + m] = ord("N")
            elif kind < 0.8:
                x = s[a:a + m]
                s[a:a + m] = np.where((x >= 65) & (x <= 90), x + 32, x)
            else:
                s[a:a + m] = np.frombuffer(b"RYKM-nx", np.uint8)[rng.integers(0, 7, len(s[a:a + m]))]
        name = f"chr{c}"
        head = [name, f"{name} description text", f"{name}x\tfield", f">{name}"][int(rng.integers(0, 4))]
        chroms.append((name, head, s.tobytes().decode()))
    genome = "".join(f">{h}\n{_wrap(s, int(rng.choice([60, 70, 13])))}\n" for _, h, s in chroms)
    names = []
    for _, h, _ in chroms:  # the names the reference's reader makes of these headers
        h = h.split(" ")[0]
        if h.startswith(">"):
            t = h.find("\t")
            names.append(h[1:] if t < 0 else h[1:t])
        else:
            t = h.find("\t")
            names.append(h if t <= 0 else h[:t - 1])

    # dump: windows of the genome (some in lower case, some with N), random k-mers, other lengths
    freq = int(rng.integers(1, 20))
    recs = []
    pool = [s for _, _, s in chroms]
    for _ in range(int(rng.integers(0, 60))):
        src = pool[int(rng.integers(0, len(pool)))]
        if len(src) >= k:
            a = int(rng.integers(0, len(src) - k + 1))
            km = src[a:a + k]
        else:
            km = "".join(rng.choice(list("ACGT"), k))
        if rng.random() < 0.3:
            km = km.lower()
        if rng.random() < 0.05:
            km = km[:max(1, k - 1)]  # another length: never matches
        recs.append((int(rng.integers(0, 3 * freq)), km))
    for _ in range(int(rng.integers(0, 10))):
        recs.append((int(rng.integers(0, 3 * freq)), "".join(rng.choice(list("ACGTN"), k))))
    rng.shuffle(recs)
    recs.append((int(rng.integers(0, 3 * freq)), "".join(rng.choice(list("ACGT"), k))))  # the last one sets k
    dump = []
    for c, km in recs:
        head = f">{c}" if rng.random() < 0.8 else f">{c} extra"
        body = km if rng.random() < 0.8 or len(km) < 2 else km[:len(km) // 2] + "\n" + km[len(km) // 2:]
        dump.append(head + "\n" + body)
    dump = "\n".join(dump) + "\n"

    # GFF3: CDS-like records (some past the chromosome end), genes with and without biotype
    gff = ["##gff-version 3"]
    for name, (_, _, s) in zip(names, chroms):
        L = len(s)
        for _ in range(int(rng.integers(0, 8))):
            a = int(rng.integers(1, L + 1))
            b = a + int(rng.integers(0, 300))
            if rng.random() < 0.85:
                b = min(b, L)
            if rng.random() < 0.2:
                a, b = b, a
            typ = ["CDS", "start_codon", "stop_codon", "exon", "gene", "gene", "mRNA"][int(rng.integers(0, 7))]
            attr = "ID=g;biotype=protein_coding" if rng.random() < 0.6 else "ID=g;biotype=lncRNA"
            gff.append("\t".join([name, "src", typ, str(a), str(b), ".", "+", ".", attr]))
        gff.append(f"{name}\tsrc\tCDS\t5")  # too few fields
    gff = "\n".join(gff) + "\n"

    # SAM of "CDS" sequences against the genome, and their fasta
    cds_recs, sam = [], ["@HD\tVN:1.6"]
    for q in range(int(rng.integers(0, 8))):
        ci = int(rng.integers(0, len(chroms)))
        L = len(chroms[ci][2])
        pos = int(rng.integers(1, L + 1))
        ops, span = [], 0
        for _ in range(int(rng.integers(1, 5))):
            op = str(rng.choice(list("MMMIDNSH")))
            n = int(rng.integers(1, 60))
            ops.append(f"{n}{op}")
        cigar = "".join(ops)
        qlen = int(rng.integers(20, 200))
        cds_recs.append((f"tx{q}", "".join(rng.choice(list("ACGT"), qlen))))
        same = int(rng.integers(0, qlen + 1))
        cs = f"cs:Z::{same // 2}*ag:{same - same // 2}" if rng.random() < 0.9 else "cs:Z:+ac"
        flag = int(rng.choice([0, 16, 256]))
        qname = f"tx{q}" if rng.random() < 0.9 else f"unknown{q}"
        sam.append("\t".join([qname, str(flag), names[ci], str(pos), "60", cigar, "*", "0", "0", "ACGT", "*",
                              f"NM:i:{int(rng.integers(0, 9))}", "ms:i:10", cs]))
    sam = "\n".join(sam) + "\n"
    cds = "".join(f">{n}\n{s}\n" for n, s in cds_recs)

    opts = {
        "width": int(rng.choice([60, 60, 7, 1, -3])),
        "soft": bool(rng.random() < 0.5),
        "alt": str(rng.choice(["n", "N", "x"])),
        "gene": bool(rng.random() < 0.3),
        "similarity": float(rng.choice([0.8, 0.3, 0.0])),
        "freq": freq,
        "use_dump": bool(rng.random() < 0.85),
        "use_gff": bool(rng.random() < 0.5),
        "use_sam": bool(rng.random() < 0.5),
    }
    return {"genome": genome, "dump": dump, "gff": gff, "sam": sam, "cds": cds, "opts": opts, "k": k,
            "names": names, "seqs": [s for _, _, s in chroms]}


def mask_args(case, d, out="out.fa"):
    """Writes a mask_case's files into folder d; returns the maskGenome flags."""
    files = {"genome.fa": case["genome"], "kmers.fa": case["dump"], "ann.gff3": case["gff"], "cds.sam": case["sam"],
             "cds.fa": case["cds"]}
    for f, t in files.items():
        with open(os.path.join(d, f), "w") as fh:
            fh.write(t)
    o = case["opts"]
    args = ["-i", "genome.fa", "-o", out, "-w", str(o["width"])]
    if o["use_dump"]:
        args += ["-k", "kmers.fa", "-f", str(o["freq"])]
    if o["use_gff"]:
        args += ["-g", "ann.gff3"] + (["-z"] if o["gene"] else [])
    if o["use_sam"]:
        args += ["-s", "cds.sam", "-c", "cds.fa", "-p", str(o["similarity"])]
    args += ["-d"] if o["soft"] else ["-l", o["alt"]]
    return args
