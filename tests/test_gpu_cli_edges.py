"""GPU: the command line's input and output edge cases, byte for byte.

Each input text goes through the drop-in CLI; the expected stdout is the
oracle's restatement of main.rs (load 286-311 with parse_allele 95-113, the
-L/-S filters 313-347 in binary64, brute-force distances, the text of
225-265), and error cases must print main.rs's message on stderr with exit
status 1.  Integer output: tolerance 0, compared as bytes.
"""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def run_cli(cgh, text, args, stdin=False):
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "in.tsv")
        with open(path, "w", encoding="utf-8", newline="") as f:
            f.write(text)
        cmd = [cgh.CLI_PATH, "-s", "-i", "-" if stdin else path, *args]
        print(f"cli {args} ({len(text)} bytes{', stdin' if stdin else ''})", flush=True)
        if stdin:
            return subprocess.run(cmd, input=text.encode(), capture_output=True, timeout=120)
        return subprocess.run(cmd, capture_output=True, timeout=120)


def expect_ok(cgh, text, args, stdin=False, **kw):
    out = run_cli(cgh, text, args, stdin)
    assert out.returncode == 0, out.stderr.decode()
    want = oracle.cli_expected(text, **kw)
    assert out.stdout.decode() == want, args


def expect_error(cgh, text, args, message):
    out = run_cli(cgh, text, args)
    assert out.returncode == 1
    assert out.stdout == b""
    assert out.stderr.decode() == f"error: {message}\n"


def mixed_text(rng, n, L, crlf=False):
    """Every kind of cell main.rs:95-113 distinguishes, ragged rows and empty
    lines."""
    kinds = ["-", "", "LNF", "PLOT3", "+5", "5.0", "0", "007", "INF-INF-5", "INF- 5", " 7 ",
             " 12 ", "\u00a012\u2003", "INF-12", "INF-", "12", "3", "3", "3", "4", "18446744073709551615",
             # Rust's trim strips White_Space only: U+001C..U+001F and U+200B stay (the cell is then not a
             # number), U+0085, U+000B, U+000C and U+3000 go
             "\x1c5", "5\x1f", "\x1e12", "\u200b3", "\x853\x85", "\x0b4\x0c", "\u300012"]
    lines = ["FILE\t" + "\t".join(f"L{j}" for j in range(L))]
    for i in range(n):
        w = L + int(rng.integers(-3, 4))
        cells = [kinds[int(k)] for k in rng.integers(0, len(kinds), max(w, 0))]
        lines.append(f"S{i}" + "".join("\t" + c for c in cells))
        if rng.random() < 0.1:
            lines.append("")
    eol = "\r\n" if crlf else "\n"
    return eol.join(lines) + eol


@pytest.mark.parametrize("crlf", [False, True])
def test_mixed_cells_ragged_rows(cgh, gpu, crlf):
    text = mixed_text(np.random.default_rng(1 + crlf), 300, 40, crlf)
    expect_ok(cgh, text, [])
    expect_ok(cgh, text, ["-r"], skip_repl=True)
    expect_ok(cgh, text, ["-t", "2", "--stripe", "5"])
    expect_ok(cgh, text, [], stdin=True)


def test_missing_value_and_separators(cgh, gpu):
    rng = np.random.default_rng(7)
    A = rng.integers(95, 101, size=(120, 30))
    text = "h," + ",".join(f"c{j}" for j in range(30)) + "\n"
    text += "".join(f"id {i}," + ",".join(str(v) for v in r) + "\n" for i, r in enumerate(A))
    # -n 99: the allele 99 reads as missing; input ',' / output '::' / index name
    args = ["-d", ",", "-n", "99", "-D", "::", "-x", "idx"]
    expect_ok(cgh, text, args, sep=",", missing="99", osep="::", index_name="idx")
    expect_ok(cgh, text, ["-d", ",", "-m", "lower-tri", "-X", "3"], sep=",", mode="lower-tri", cap=3)
    # a bare CR ending the text is a line of its own (a sample named "\r")
    expect_ok(cgh, text + "\r", ["-d", ","], sep=",")


def test_completeness_filters(cgh, gpu):
    n, L = 100, 12
    rng = np.random.default_rng(3)
    A = rng.integers(1, 6, size=(n, L))
    A[57:, 0] = 0                       # 57% called: dropped by -L 57 in binary64
    A[rng.random((n, L)) < 0.2] = 0
    A[:, 0] = np.where(np.arange(n) < 57, np.maximum(A[:, 0], 1), 0)
    text = "FILE\t" + "\t".join(f"L{j}" for j in range(L)) + "\n"
    text += "".join(f"S{i}\t" + "\t".join("-" if v == 0 else str(v) for v in r) + "\n" for i, r in enumerate(A))
    for lc, sc in ((57, None), (56.5, None), (None, 80), (75, 90), (100, 100), (0, 0)):
        args = (["-L", str(lc)] if lc is not None else []) + (["-S", str(sc)] if sc is not None else [])
        expect_ok(cgh, text, args, locus_pct=lc, sample_pct=sc)


def test_degenerate_shapes(cgh, gpu):
    expect_ok(cgh, "FILE\tA\tB\nonly\t1\t2\n", [])                 # n = 1
    expect_ok(cgh, "FILE\nS0\nS1\t5\nS2\n", [])                     # L = 0
    expect_ok(cgh, "FILE\tA\tB\n", [])                              # n = 0
    expect_ok(cgh, "FILE\tA\tB\n\n\n", ["-m", "upper-tri"])         # n = 0
    expect_ok(cgh, "\n", [])                                        # empty header line
    expect_ok(cgh, "FILE\tA\nS0\t-\nS1\tLNF\nS2\t\n", [])           # everything missing
    expect_ok(cgh, "FILE\tA\tB\nS0\t1\t2\nS1\t1\t3\n", ["-L", "100", "-S", "100", "-X", "1"],
              locus_pct=100, sample_pct=100, cap=1)
    expect_ok(cgh, "FILE\tA\nS0\t1\nS1\t-\n", ["-L", "60"], locus_pct=60)  # every locus dropped


def test_input_errors(cgh, gpu):
    expect_error(cgh, "", [], "empty input")
    big = "18446744073709551616"
    expect_error(cgh, f"FILE\tA\tB\nS0\t1\t2\nS1\t{big}\t2\n", [], f"sample S1: allele '{big}' too large")
    expect_error(cgh, f"FILE\tA\nS0\t1\nS1\t2\t INF-{big} \n", [], f"sample S1: allele 'INF-{big}' too large")
    # 65,536 loci do not fit u16 distances
    L = 65536
    expect_error(cgh, "FILE" + "\tx" * L + "\nS0" + "\t1" * L + "\nS1" + "\t2" * L + "\n", [],
                 "more than 65535 loci (u16 distances)")
    # 65,536 distinct alleles at one locus
    n = 65536
    text = "FILE\tA\tB\n" + "".join(f"S{i}\t{i + 1}\t1\n" for i in range(n))
    expect_error(cgh, text, [], "locus 0 has 65536 distinct alleles (limit 65535)")
