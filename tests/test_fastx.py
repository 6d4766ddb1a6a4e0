"""FASTA / FASTQ ingestion (include/pa_fastx.h): host-side parsing into residues + offsets.

Checked against the reference's FASTQ fixture (tests/golden/in.fastq is test/data/in.fastq,
SHA-256 6e30fc35... as declared in test/data/datasources.cmake:6-8), and against a plain
sequential parser written here from the format rules on seeded random files: multi-line FASTA,
CRLF line ends, blank and ';' comment lines, empty records, quality lines that start with '@' or
'+', lowercase residues with and without folding, and inputs of several MiB so that the parallel
chunking cuts between records.  The GPU test scores pairs read from files through the C ABI and
compares with the oracle.
"""
import hashlib
import os

import numpy as np
import pytest

from pairwise_aligner_amd import device, fastx, workloads

HERE = os.path.dirname(os.path.abspath(__file__))
IN_FASTQ = os.path.join(HERE, "golden", "in.fastq")


def simple_parse(text: bytes, uppercase=False):
    """Sequential restatement of the format rules: (names, sequences, qualities or None)."""
    lines = text.split(b"\n")
    first = text.lstrip(b" \t\r\n")[:1]
    fold = (lambda s: s.upper()) if uppercase else (lambda s: s)
    names, seqs, quals = [], [], []
    if first == b">":
        for ln in lines:
            if ln.startswith(b">"):
                names.append(ln[1:].rstrip(b"\r"))
                seqs.append(b"")
            elif ln.startswith(b";") or not ln.strip(b" \t\r"):
                continue
            else:
                seqs[-1] += fold(ln.replace(b" ", b"").replace(b"\t", b"").replace(b"\r", b""))
        return names, seqs, None
    i = 0
    while i < len(lines):
        if not lines[i].strip(b" \t\r"):
            i += 1
            continue
        assert lines[i].startswith(b"@")
        names.append(lines[i][1:].rstrip(b"\r"))
        seqs.append(fold(lines[i + 1].rstrip(b"\r")))
        assert lines[i + 2].startswith(b"+")
        quals.append(lines[i + 3].rstrip(b"\r") if i + 3 < len(lines) else b"")
        i += 4
    return names, seqs, quals


def check_against_simple(text: bytes, uppercase=False):
    recs = fastx.parse(text, uppercase=uppercase)
    names, seqs, quals = simple_parse(text, uppercase)
    assert len(recs) == len(seqs)
    assert recs.names == [n.decode("latin-1") for n in names]
    assert recs.sequences() == seqs
    assert recs.offsets[0] == 0 and np.all(np.diff(recs.offsets.astype(np.int64)) >= 0)
    if quals is None:
        assert recs.format == fastx.FASTA and recs.qualities is None
    else:
        assert recs.format == fastx.FASTQ
        got_q = [recs.qualities[recs.offsets[r]:recs.offsets[r + 1]].tobytes() for r in range(len(recs))]
        assert got_q == quals
    return recs


def test_reference_fixture():
    data = open(IN_FASTQ, "rb").read()
    assert hashlib.sha256(data).hexdigest().startswith("6e30fc35f908a36fe0c68a7a35c47f51f9570da16622fb0c")
    recs = fastx.read(IN_FASTQ)
    assert recs.format == fastx.FASTQ
    assert recs.names == ["seq1", "seq2"]
    assert recs.sequences() == [b"ACGTTTGATTCGCG", b"TCGGGGGATTCGCG"]
    assert recs.qualities.tobytes() == b"!**55CCF>>>>>C" + b"!**65CCF>CF>>C"
    check_against_simple(data)


def random_fasta(rng, n, max_len, crlf=False, lower=False, width=60):
    out = []
    nl = b"\r\n" if crlf else b"\n"
    for r in range(n):
        L = int(rng.integers(0, max_len + 1))
        s = bytes(rng.choice(list(b"ACGTN" if not lower else b"acgtnACGT"), size=L).tolist())
        out.append(b">rec%d some description" % r + nl)
        if r % 17 == 3:
            out.append(b";comment line" + nl)
        for k in range(0, L, width):
            out.append(s[k:k + width] + nl)
        if r % 11 == 5:
            out.append(nl)
    return b"".join(out)


def random_fastq(rng, n, max_len, crlf=False):
    out = []
    nl = b"\r\n" if crlf else b"\n"
    for r in range(n):
        L = int(rng.integers(0, max_len + 1))
        s = bytes(rng.choice(list(b"ACGT"), size=L).tolist())
        q = bytes(rng.integers(33, 75, size=L).astype(np.uint8).tolist())
        if L and r % 5 == 0:
            q = b"@" + q[1:]        # quality lines that look like headers
        if L and r % 7 == 0:
            q = b"+" + q[1:]
        plus = b"+" if r % 2 else b"+read%d" % r
        out.append(b"@read%d" % r + nl + s + nl + plus + nl + q + nl)
        if r % 13 == 0:
            out.append(nl)
    return b"".join(out)


@pytest.mark.parametrize("crlf", [False, True])
def test_fasta_small_and_multiline(crlf):
    rng = np.random.default_rng(11 + crlf)
    check_against_simple(random_fasta(rng, 300, 400, crlf=crlf))
    check_against_simple(random_fasta(rng, 50, 200, crlf=crlf, width=7))


def test_fasta_lowercase_folding():
    rng = np.random.default_rng(13)
    text = random_fasta(rng, 200, 300, lower=True)
    a = check_against_simple(text, uppercase=False)
    b = check_against_simple(text, uppercase=True)
    assert any(c in a.residues for c in b"acgt")
    assert not any(c in b.residues for c in b"acgtn")


@pytest.mark.parametrize("crlf", [False, True])
def test_fastq_small(crlf):
    rng = np.random.default_rng(17 + crlf)
    check_against_simple(random_fastq(rng, 500, 300, crlf=crlf))


def test_large_inputs_cut_between_records():
    """Several MiB: the parser cuts the text into chunks at record starts on many threads."""
    rng = np.random.default_rng(19)
    check_against_simple(random_fasta(rng, 40000, 500))
    check_against_simple(random_fastq(rng, 30000, 400))


def test_empty_and_blank_inputs():
    for text in (b"", b"\n\n  \n"):
        recs = fastx.parse(text)
        assert len(recs) == 0 and len(recs.residues) == 0
    recs = fastx.parse(b">a\n>b\nAC\n>c\n")
    assert recs.sequences() == [b"", b"AC", b""]
    recs = fastx.parse(b"@a\n\n+\n\n@b\nA\n+\nI")   # empty record, no final newline
    assert recs.sequences() == [b"", b"A"] and recs.qualities.tobytes() == b"I"
    recs = fastx.parse(b">x\r\nAC GT\r\nTT\r")   # CRLF, a blank inside a line, no final newline
    assert recs.names == ["x"] and recs.sequences() == [b"ACGTTT"]
    recs = fastx.parse(b">only header")
    assert recs.names == ["only header"] and recs.sequences() == [b""]


@pytest.mark.parametrize("text, what", [
    (b"ACGT\n", "neither FASTA"),
    (b"@r\nACGT\n+\nIII\n", "quality line length"),
    (b"@r\nACGT\nACGT\n+\nIIIIIIII\n", "'+' line"),
    (b"@r\n", "sequence line"),
])
def test_malformed_inputs_raise(text, what):
    with pytest.raises(device.PaError) as e:
        fastx.parse(text)
    assert what in str(e.value)


def test_arrays_outlive_the_records():
    """The arrays are views of the library's buffers; the buffers live while any view does."""
    import gc
    rng = np.random.default_rng(31)
    text = random_fastq(rng, 2000, 200)
    want = simple_parse(text)[1]
    recs = fastx.parse(text)
    res, off = recs.residues, recs.offsets
    del recs
    gc.collect()
    for _ in range(3):
        fastx.parse(random_fasta(rng, 500, 300))   # reuse of freed host memory would show here
    assert [res[off[r]:off[r + 1]].tobytes() for r in range(len(off) - 1)] == want


def test_missing_file_raises(tmp_path):
    with pytest.raises(device.PaError):
        fastx.read(str(tmp_path / "absent.fa"))


def test_file_round_trip(tmp_path):
    rng = np.random.default_rng(23)
    s = workloads.protein_pairs(rng, 200, 0, 300)[0]
    path = tmp_path / "prot.fa"
    path.write_bytes(b"".join(b">p%d\n" % i + x + b"\n" for i, x in enumerate(s)))
    recs = fastx.read(str(path))
    assert recs.sequences() == list(s)
    b, o = workloads.pack(s)
    assert np.array_equal(recs.residues, b) and np.array_equal(recs.offsets, o)


@pytest.mark.gpu
def test_scores_of_pairs_read_from_files(tmp_path):
    from oracle import oracle as O
    rng = np.random.default_rng(29)
    s1, s2 = workloads.dna_pairs(rng, 2000, 10, 400, divergence=0.1, indel_fraction=0.3)
    (tmp_path / "a.fq").write_bytes(b"".join(b"@a%d\n%s\n+\n%s\n" % (i, x, b"I" * len(x)) for i, x in enumerate(s1)))
    (tmp_path / "b.fa").write_bytes(b"".join(b">b%d\n%s\n" % (i, x.lower()) for i, x in enumerate(s2)))
    a = fastx.read(str(tmp_path / "a.fq"))
    b = fastx.read(str(tmp_path / "b.fa"), uppercase=True)
    dc = device.DeviceConfig(match=4, mismatch=-5, gap_open=-10, gap_extension=-1)
    ctx = device.Context(0)
    got, _ = ctx.score_packed(a.residues, a.offsets, b.residues, b.offsets, dc)
    want = O.score_packed(a.residues, a.offsets, b.residues, b.offsets,
                          O.OracleConfig(gap_open=-10, gap_extension=-1, match=4, mismatch=-5), threads=8)
    assert np.array_equal(got.astype(np.int64), want)
