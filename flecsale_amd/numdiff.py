"""Column-by-column numerical comparison of two text files: the rule the
reference's regression tests apply to a dump and its .std golden
(tools/numdiff.py, driven by the create_regression_test COMPARE/STANDARD
pairs of apps/*/CMakeLists.txt).

Lines are compared in order after trailing white space is stripped; the
comparison stops at the first empty line of the expected file (the actual one
must end there too).  Each line is split on commas, semicolons and runs of
white space (a trailing empty field is dropped); the two lines must have as
many fields.  A pair of fields that both read as numbers fails when
|actual - expected| exceeds |expected| * rel_tol (if rel_tol > 0) or abs_tol
(if abs_tol > 0); with neither given, abs_tol = tol (1e-15).  Other fields
are ignored unless check_text, then they must match exactly.

    python -m flecsale_amd.numdiff [-a ABS] [-r REL] [-t TOL] [-c] expected actual
"""
import argparse
import re
import sys

_FIELDS = re.compile(r",|\s+|;")


def _fields(line):
    parts = _FIELDS.split(line)
    if parts and parts[-1] == "":
        parts.pop()
    return parts


def equivalent(expected, actual, rel_tol, abs_tol):
    err = abs(actual - expected)
    if rel_tol > 0 and err > abs(expected) * rel_tol:
        return False
    if abs_tol > 0 and err > abs_tol:
        return False
    return True


def compare(expected_path, actual_path, rel_tol=-1.0, abs_tol=-1.0, tol=1e-15, check_text=False):
    """The failures (human-readable strings) of `actual_path` against
    `expected_path`; an empty list means the files agree."""
    if rel_tol < 0 and abs_tol < 0:
        abs_tol = tol
    failures = []
    with open(expected_path) as fe, open(actual_path) as fa:
        line = 0
        while True:
            line += 1
            exp = fe.readline().rstrip()
            act = fa.readline().rstrip()
            if exp == "":
                if act != "":
                    failures.append("tested file ended too late")
                break
            if act == "":
                failures.append("tested file ended too early")
                break
            ef, af = _fields(exp), _fields(act)
            if len(ef) != len(af):
                failures.append(f"wrong number of columns in line {line}")
                continue
            for col, (e, a) in enumerate(zip(ef, af)):
                try:
                    ev, av = float(e), float(a)
                except ValueError:
                    if check_text and e != a:
                        failures.append(f"text did not match in line {line}, column {col}")
                    continue
                if not equivalent(ev, av, rel_tol, abs_tol):
                    failures.append(f"non-equivalence in line {line}, column {col}: "
                                    f"expected {ev:.15e}, got {av:.15e}")
    return failures


def main(argv=None):
    ap = argparse.ArgumentParser(description="numerical diff of a dump against its standard")
    ap.add_argument("expected")
    ap.add_argument("actual")
    ap.add_argument("-t", "--tolerance", type=float, default=1e-15, dest="tol")
    ap.add_argument("-r", "--relative", type=float, default=-1.0, dest="rel_tol")
    ap.add_argument("-a", "--absolute", type=float, default=-1.0, dest="abs_tol")
    ap.add_argument("-c", "--check-text", action="store_true")
    a = ap.parse_args(argv)
    failures = compare(a.expected, a.actual, a.rel_tol, a.abs_tol, a.tol, a.check_text)
    for f in failures[:20]:
        print(">>>> " + f)
    print("FAILURE" if failures else "SUCCESS")
    return 1 if failures else 0


if __name__ == "__main__":
    sys.exit(main())
