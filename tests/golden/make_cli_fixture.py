"""Writes tests/golden/cli_behaviour.json: exit status, stdout and stderr of
the ratpoints-compatible command line for argument lists that probe its
parsing rules (option table, decimal reals, intervals, device lists, batch
sizes, heights, coefficient lists) and its output formatter.

The expected bytes are the command line's interface (the messages, order of
checks and output format of src/command_line.cpp, src/output_formatter.cpp
and src/main.cpp in the reference).  They were captured from this
repository's CLI built over the oracle's CPU test double of the C ABI
(oracle/_build/ratpoints_gpu_oracle, host logic only) while its parser still
followed the reference's control flow statement by statement, before the
parser, formatter and driver were rewritten; tests/test_cli.py holds the
rewritten ones to the same bytes.

Usage: python tests/golden/make_cli_fixture.py [BINARY]
"""

import json
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def cases():
    c = "1 0 1"
    out = []
    # decimal reals as interval ends
    for v in ["0", "-0", "+0", "1", "-1.5", "1.", ".5", "-.5", ".", "+", "-", "", "e5", "1e", "1e+",
              "1e-", "1e5", "1E-3", "2.5e-1", "1e10000", "1e-10000", "1e10001", "1e99999x",
              "1e0010000", "1e00010001", "12x", "x12", "1.2.3", "1e5.0", "--1", "+-1", "0x10",
              " 1", "1 ", "1/2", "3.14159", "-7e+2", "00012.5000", "1e+0", "9" * 40]:
        out.append([c, "30", "-du", "6", "-l", v])
        out.append([c, "30", "-du", "6", "-u", v])
    # interval sequences
    seqs = [["-l", "1", "-u", "2"], ["-l", "2", "-u", "1"], ["-l", "1", "-l", "2"],
            ["-u", "1", "-u", "2"], ["-u", "2", "-u", "1"], ["-u", "1", "-l", "2", "-u", "3"],
            ["-l", "0", "-u", "1", "-l", "1", "-u", "2"], ["-l", "0", "-u", "1", "-l", "0.5", "-u", "2"],
            ["-l", "0", "-u", "1", "-u", "2"], ["-l", "0"], ["-l", "0", "-u", "1", "-l", "3"],
            ["-l", "1", "-l"], ["-l"], ["-u"], ["-l", "3", "-u", "3"], ["-u", "-1", "-l", "-0.5", "-u", "0"],
            ["-l", "-1", "-u", "1", "-l", "2"], ["-l", "5", "-u", "6", "-l", "1"]]
    for s in seqs:
        out.append([c, "30", "-du", "6", *s])
    # devices, batch sizes, denominators, heights
    for v in ["all", "0", "00", "0,", ",0", "0,,1", "0,0", "1,0", "-1", "+0", "a", "0 ", "all,0",
              "99999999999999999999", "2147483647", "2147483648", "7", "0,1", ","]:
        out.append([c, "5", "--devices", v])
    for v in ["1", "0", "-1", "65536", "65537", "abc", "3.5", "+7", " 8", "8 ", "99999999999999999999"]:
        out.append([c, "5", "--batch-size", v])
    for opt in ["-dl", "-du"]:
        for v in ["1", "0", "-3", "x", "2147483647", "2147483648", "99999999999999999999", "6", "3"]:
            out.append([c, "5", opt, v])
    for h in ["0", "-5", "5x", "x", " 5", "+5", "4611686018427387887", "4611686018427387888",
              "99999999999999999999", "2147483648"]:
        out.append([c, h, "-du", "3", "-z"])
    # coefficient lists
    for co in ["", "1", "5", "1 0", "0 1", "1 0 0", "1 -0", "1 +0", "+1 0 +1", "++1 0 1", "1 a",
               "1 1.5", "1\t0\n1", "  1   0   1  ", "-", "1 -", "1 " + "0 " * 99 + "1",
               "1 " + "0 " * 100 + "1", "1 -2 1", "0 0 1 1"]:
        out.append([co, "4", "-du", "3"])
    # unknown and flag options, option needing a value at the end
    for o in [["-x"], ["--devices"], ["--batch-size"], ["-dl"], ["-f"], ["-fs"], ["-s", "-k", "-K", "-j", "-J"],
              ["-q", "-v"], ["--square-denominators", "-du", "99999"], ["--help"], ["-1", "-1"], ["-"]]:
        out.append([c, "12", *o])
    # output formats
    for f in ["%x", "%y", "%z", "%q", "%", "%%", "100\\%", "\\%x", "\\", "a\\", "\\n\\t\\\\\\x",
              "%x %y %z\\n", "[%y]", "%\\n", "%\\%", "%y%y%x%z\\t\\n"]:
        out.append(["1 0 1", "12", "-f", f])
        out.append(["1 0 1", "12", "-y", "-f", f])
    for extra in [["-z", "-fs", "<", "-fm", "|", "-fe", ">"], ["-fs", "<", "-fm", "|", "-fe", ">\\n"],
                  ["-1", "-fm", "|"], ["-i", "-1"], ["-Z", "-z"], ["-z", "-Z"], ["-y", "-Y"], ["-I", "-i"]]:
        out.append(["1 0 1", "12", *extra])
    # escapes in the texts around the points ("\%" stays the byte 0x01 there)
    for extra in [["-fs", "a\\%b", "-fm", "\\%", "-fe", "\\%\\n"], ["-z", "-fe", "end\\%"],
                  ["-f", "%x\x01%z\\n"], ["-f", "%\\t%\\x%"], ["-fs", "\\", "-fe", "x\\"]]:
        out.append(["1 0 1", "12", *extra])
    out += [[], ["--help"], ["--help", "x"], ["1 0 1"], ["--list-devices", "x"]]
    return out


def main(argv):
    binary = argv[1] if len(argv) > 1 else os.path.join(ROOT, "oracle", "_build",
                                                        "ratpoints_gpu_oracle")
    env = dict(os.environ)
    env.pop("RATPOINTS_GPU_BENCHMARK", None)
    rows = []
    for args in cases():
        p = subprocess.run([binary, *args], capture_output=True, timeout=60, env=env)
        rows.append({"args": args, "rc": p.returncode, "stdout": p.stdout.decode("latin-1"),
                     "stderr": p.stderr.decode("latin-1")})
    with open(os.path.join(HERE, "cli_behaviour.json"), "w") as fh:
        json.dump(rows, fh, indent=0)
    print(f"wrote {len(rows)} cases")


if __name__ == "__main__":
    main(sys.argv)
