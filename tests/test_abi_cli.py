"""CPU-side checks of the boundary: libxtn.so loads and exports every
symbol include/xtn.h declares, and the ctypes mirrors of its structs have
the header's layout; the xt_neighbor CLI's argument and input
parsing paths (which finish before the GPU is touched) reproduce the
reference's messages and exit codes (cli.cu:29-193, airr.cu:116-151)."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import xtn
from conftest import GOLDEN, REPO

HEADER = os.path.join(REPO, "include", "xtn.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(xtn_[a-z_]+)\s*\(", text)) - {"xtn_callback"})


def test_library_exports_declared_symbols():
    lib = xtn.load_library()
    names = declared_functions()
    assert set(names) == set(xtn.EXPORTED)
    for name in names:
        assert hasattr(lib, name), name


def test_version_status_and_defaults():
    lib = xtn.load_library()
    assert lib.xtn_version() == b"2.0.0-mi355x"
    assert lib.xtn_status_string(0) == b"ok"
    assert lib.xtn_status_string(6) == b"capacity exceeded"
    cfg = xtn.Config()
    lib.xtn_config_init(ctypes.byref(cfg))
    assert (cfg.distance, cfg.measure, cfg.device, cfg.max_records_per_chunk) == (1, 0, 0, 0)


def test_struct_layouts_match_header(tmp_path):
    # every struct the bindings mirror has the header's size and field offsets
    # (a C program built against include/xtn.h prints them)
    structs = {"xtn_int2": xtn.Int2, "xtn_seq_info": xtn.SeqInfo, "xtn_output": xtn.Output,
               "xtn_config": xtn.Config, "xtn_server_config": xtn.ServerConfig, "xtn_stats": xtn.Stats,
               "xtn_tuning": xtn.Tuning}
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "xtn.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(REPO, "include"), "-o", str(exe), str(src)],
                   check=True, timeout=60)
    got = dict(line.split() for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                       check=True, timeout=60).stdout.splitlines())
    for cname, py in structs.items():
        assert int(got[cname]) == ctypes.sizeof(py), cname
        for f, _ in py._fields_:
            assert int(got[f"{cname}.{f}"]) == getattr(py, f).offset, (cname, f)


def test_open_rejects_bad_config():
    lib = xtn.load_library()
    cfg = xtn.Config()
    lib.xtn_config_init(ctypes.byref(cfg))
    cfg.distance = 0
    h = ctypes.c_void_p()
    assert lib.xtn_open(ctypes.byref(h), ctypes.byref(cfg)) == 3
    cfg.distance = 1
    cfg.measure = 5
    assert lib.xtn_open(ctypes.byref(h), ctypes.byref(cfg)) == 3


def run_cli(args, **kw):
    env = dict(os.environ, LC_ALL="C")
    return subprocess.run([xtn.CLI_PATH] + args, capture_output=True, env=env, timeout=60, **kw)


def write(tmp_path, name, text):
    p = tmp_path / name
    p.write_bytes(text.encode() if isinstance(text, str) else text)
    return str(p)


def test_cli_version_help():
    r = run_cli(["-v"])
    assert (r.returncode, r.stdout, r.stderr) == (2, b"2.0.0\n", b"")
    r = run_cli(["--help"])
    assert r.returncode == 2
    assert r.stdout.startswith(b"xt_neighbor: perform either nearest neighbor search")
    assert b"-N or --info-length [number] (required)" in r.stdout
    # -v wins over a later unknown flag, as flags are read left to right
    assert run_cli(["-v", "--bogus"]).returncode == 2


@pytest.mark.parametrize("args,msg", [
    (["--bogus"], "unknown option"),
    (["-n", "5"], "missing path for seq"),
    (["-i", "x"], "missing length for seq"),
    (["-i", "x", "-n", "0"], "invalid input length"),
    (["-i", "x", "-n", "3", "-d", "0"], "distance must be a valid number larger than 0"),
    (["-i", "x", "-n", "3", "-d", "-2"], "distance must be a valid number larger than 0"),
    (["-i", "x", "-n", "3", "-m", "lev"], "invalid measure option"),
    (["-i", "x", "-n", "3", "-N", "0"], "invalid info length"),
    (["-i", "x", "-n", "3", "-I", "info.csv"], "repertiore path or repertoire count is missing in overlap mode"),
    (["-i", "x", "-n", "3", "-N", "2"], "repertiore path or repertoire count is missing in overlap mode"),
    (["-i", "/nonexistent/file", "-n", "3"], "input file reading failed"),
])
def test_cli_argument_errors(args, msg):
    r = run_cli(args)
    assert r.returncode == 1
    assert r.stderr == f"Error: {msg}\n".encode()
    assert r.stdout == b""


def test_cli_invalid_residue():
    path = os.path.join(GOLDEN, "sample_input3.txt")  # CAAA / AAAA / AAAX, header skipped
    r = run_cli(["-i", path, "-n", "2"])
    assert r.returncode == 1
    assert r.stderr == b"Error: invalid char\nError: input parsing error\nError line: 3\n"


def test_cli_length_mismatch(tmp_path):
    p = write(tmp_path, "in.csv", "cdr3\nCAAA\nCADA\n\n \nCDKD\n")
    r = run_cli(["-i", p, "-n", "4"])
    assert r.returncode == 1
    assert r.stderr == b"Error: input length doesn't match with the actual\n"
    # more rows than -n overflow the sequence store first
    r = run_cli(["-i", p, "-n", "2"])
    assert r.stderr == b"Error: SeqArray exceeding max length\nError: input parsing error\nError line: 6\n"


def test_cli_crlf_rejected(tmp_path):
    p = write(tmp_path, "in.csv", "cdr3\r\nCAAA\r\n")
    r = run_cli(["-i", p, "-n", "1"])
    assert r.stderr == b"Error: invalid char\nError: input parsing error\nError line: 2\n"


def test_cli_overlap_parse_errors(tmp_path):
    seqs = write(tmp_path, "seqs.csv", "cdr3,count\nCAAA,3\nCADA,x\n")
    info = write(tmp_path, "info.csv", "size\n1\n1\n")
    r = run_cli(["-i", seqs, "-n", "2", "-I", info, "-N", "2"])
    assert r.stderr == b"Error: input parsing error (invalid number)\nError line: 3\n"
    seqs2 = write(tmp_path, "seqs2.csv", "cdr3,count\nCAAA,3\nCADA\n")
    r = run_cli(["-i", seqs2, "-n", "2", "-I", info, "-N", "2"])
    assert r.stderr == b"Error: input parsing error (comma expected)\nError line: 3\n"
    bad_info = write(tmp_path, "bad.csv", "size\n1\n0\n")
    r = run_cli(["-i", seqs, "-n", "2", "-I", bad_info, "-N", "2"])
    assert r.stderr == b"Error: info parsing error (invalid number)\nError line: 3\n"
    big_info = write(tmp_path, "big.csv", "size\n2\n1\n")
    r = run_cli(["-i", seqs, "-n", "2", "-I", big_info, "-N", "2"])
    assert r.stderr == b"Error: total repertoires' size does not match sequence count [1]\n"
    short_info = write(tmp_path, "short.csv", "size\n1\n")
    r = run_cli(["-i", seqs, "-n", "2", "-I", short_info, "-N", "2"])
    assert r.stderr == b"Error: info length doesn't match with the actual\n"
    small_info = write(tmp_path, "small.csv", "size\n1\n")
    r = run_cli(["-i", seqs, "-n", "2", "-I", small_info, "-N", "1"])
    assert r.stderr == b"Error: total repertoires' size does not match sequence count [2]\n"
    r = run_cli(["-i", seqs, "-n", "2", "-I", "/nonexistent", "-N", "1"])
    assert r.stderr == b"Error: info file reading failed\n"


def test_cli_airr_parse_errors(tmp_path):
    hdr = "a\tduplicate_count\tcdr3_aa\n"
    bad_seq = write(tmp_path, "a1.tsv", hdr + "x\t1\tCAAA\ny\t2\tCAXA\n")
    r = run_cli(["-a", "-i", bad_seq, "-n", "2"])
    # data rows are numbered from 0 in AIRR mode (airr.cu:122, 130)
    assert r.stderr == (b"Error: invalid char\nError: invalid seq\nError line: 1\n"
                        b"Error: line parsing error\nError line: 1\n")
    short = write(tmp_path, "a2.tsv", hdr + "x\t1\n")
    r = run_cli(["-a", "-i", short, "-n", "1"])
    assert r.stderr == b"Error: line parsing error\nError line: 0\n"
    info = write(tmp_path, "info.csv", "size\n2\n")
    bad_num = write(tmp_path, "a3.tsv", hdr + "x\t1\tCAAA\ny\t0\tCADA\n")
    r = run_cli(["-a", "-i", bad_num, "-n", "2", "-I", info, "-N", "1"])
    assert r.stderr == b"Error: invalid number\nError line: 1\nError: line parsing error\nError line: 1\n"


def _dump(tmp_path, args, serial):
    env = dict(os.environ)
    dump = str(tmp_path / ("serial.bin" if serial else "parallel.bin"))
    env["XTN_CLI_PARSE_DUMP"] = dump
    env.pop("XTN_CLI_SERIAL_PARSE", None)
    if serial:
        env["XTN_CLI_SERIAL_PARSE"] = "1"
    r = subprocess.run([xtn.CLI_PATH] + args, capture_output=True, env=env)
    data = open(dump, "rb").read() if os.path.exists(dump) and r.returncode == 0 else b""
    if os.path.exists(dump):
        os.remove(dump)
    return r.returncode, r.stdout, r.stderr, data


@pytest.mark.parametrize("case", ["plain", "blank_lines", "two_columns", "no_final_newline", "long_line",
                                  "bad_char_late", "freq_on_next_line", "count_mismatch"])
def test_parallel_parser_equals_serial(tmp_path, case):
    """the threaded CSV parser (files over 1 MB) gives the serial parser's
    store, output and messages byte for byte, including the cases it hands
    back to the serial parser"""
    rng = np.random.default_rng(sum(map(ord, case)))
    seqs = ["".join(rng.choice(list("ACDEFGHIKLMNPQRSTVWY"), int(rng.integers(6, 22)))) for _ in range(80000)]
    two = case == "two_columns" or case == "freq_on_next_line"
    lines = [s + (f",{int(rng.integers(1, 99999))}" if two else "") for s in seqs]
    n = len(lines)
    if case == "blank_lines":
        for k in range(0, n, 997):
            lines.insert(k, "" if k % 2 else " ")
    if case == "long_line":
        lines[40000] = "C" * 600
    if case == "bad_char_late":
        lines[79000] = "CASSXF"
    if case == "freq_on_next_line":
        lines[50000] = seqs[50000] + ",  "
    body = "cdr3" + (",freq" if two else "") + "\n" + "\n".join(lines)
    if case != "no_final_newline":
        body += "\n"
    path = tmp_path / "in.csv"
    path.write_text(body)
    count = n + (1 if case == "count_mismatch" else 0)
    args = ["-i", str(path), "-n", str(count), "-d", "1"]
    if two:
        info = tmp_path / "info.csv"
        info.write_text(f"size\n{n // 2}\n{n - n // 2}\n")
        args += ["-I", str(info), "-N", "2"]
    ser = _dump(tmp_path, args, True)
    par = _dump(tmp_path, args, False)
    assert par == ser
    if case in ("plain", "blank_lines", "two_columns", "no_final_newline"):
        assert ser[0] == 0 and len(ser[3]) > 4 * n


def run_cli1(args, **kw):
    return subprocess.run([xtn.CLI1_PATH] + args, capture_output=True, **kw)


def test_cli1_flags_and_errors(tmp_path):
    """the non-streaming xtneighbor command line (xtneighbor/src/cli.cu):
    version, help, argument and input errors, all before the device"""
    r = run_cli1(["-v"])
    assert r.returncode == 2 and r.stdout == b"1.0.0\n"
    r = run_cli1(["-h"])
    assert r.returncode == 2 and r.stdout.startswith(b"xt_neighbor\n\t description: perform xt_neighbor")
    assert b"-c or --check-output" in r.stdout
    cases = [
        (["-n", "3"], "missing path for seq1"),
        (["-p", "x"], "missing length for seq1"),
        (["-p", "x", "-n", "3", "-d", "0"], "distance must be a valid number larger than 0"),
        (["-p", "x", "-n", "abc"], "invalid input length"),
        (["-p", "x", "-n", "3", "-m", "leven"], "unknown option"),
        (["-p", str(tmp_path / "missing.txt"), "-n", "3"], "file reading failed"),
    ]
    for args, msg in cases:
        r = run_cli1(args)
        assert r.returncode == 1 and r.stderr == f"Error: {msg}\n".encode(), args
    r = run_cli1(["-p", os.path.join(GOLDEN, "sample_input3.txt"), "-n", "3"])
    assert r.returncode == 1
    assert r.stderr == (b"Error: parsing error at line: 3 (only upper-cased amino acids with max length of 18 "
                        b"are allowed)\n")
    r = run_cli1(["-p", os.path.join(GOLDEN, "sample_input1.txt"), "-n", "5"])
    assert r.returncode == 1 and r.stderr == b"Error: input length doesn't match with the actual\n"
    # a 49-character line is read in two fgets calls, the second "\n" skipped
    p = tmp_path / "long.txt"
    p.write_text("C" * 49 + "\nCASS\n")
    r = run_cli1(["-p", str(p), "-n", "3"])
    assert r.returncode == 1 and r.stderr == b"Error: input length doesn't match with the actual\n"
