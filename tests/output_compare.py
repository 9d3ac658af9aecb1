"""Comparison of a text output file with the reference's own: the same lines,
the same tokens and the same layout, numbers allowed to differ in their last
printed digits.

A line that differs from the reference's must still have the same shape once
every number is replaced by the pattern of its printed form (sign, digits
before and after the point, exponent width), so widths, precisions and
separators are pinned exactly; the numbers of such a line must agree within
rtol (relative to the larger magnitude) or atol.
"""
import re

NUM = re.compile(r"[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[eE][-+]?\d+)?")


def _pattern(tok):
    m = re.fullmatch(r"([-+]?)(\d*)(\.?)(\d*)(?:([eE])([-+]?)(\d+))?", tok)
    sign, ip, dot, fp, e, es, ed = m.groups()
    if e:  # scientific: mantissa digits and exponent width, not the values
        return f"<{sign}{len(ip)}{dot}{len(fp)}e{len(ed)}>"
    return f"<{sign}{dot}num>"  # fixed or %g: its width follows the value


def shape(line):
    return NUM.sub(lambda m: _pattern(m.group()), line)


def compare(ref_text, got_text, rtol=1e-8, atol=1e-12, name=""):
    """Returns the number of lines that differ in text (all within tolerance);
    raises AssertionError naming the first line that breaks the layout or the
    tolerance."""
    ref, got = ref_text.splitlines(), got_text.splitlines()
    assert len(ref) == len(got), f"{name}: {len(got)} lines, the reference has {len(ref)}"
    differing = 0
    for k, (a, b) in enumerate(zip(ref, got)):
        if a == b:
            continue
        differing += 1
        assert shape(a) == shape(b), f"{name}:{k + 1}: layout differs\n  ref: {a!r}\n  got: {b!r}"
        ta, tb = NUM.findall(a), NUM.findall(b)
        assert NUM.sub("#", a) == NUM.sub("#", b), f"{name}:{k + 1}: text differs"
        for x, y in zip(ta, tb):
            fx, fy = float(x), float(y)
            assert abs(fx - fy) <= max(atol, rtol * max(abs(fx), abs(fy))), \
                f"{name}:{k + 1}: {y} against the reference's {x}: beyond rtol {rtol}"
    return differing
