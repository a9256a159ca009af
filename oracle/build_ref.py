"""Compile the reference's own CPU C kernels into ``oracle/_ref/``.

TEST INFRASTRUCTURE ONLY.  The reference keeps its CPU kernels as C source
held in Python string literals and compiles them at run time with cffi.  This
recipe reads ``Beam.py`` as text where it lies under /root/reference, takes
the two string literals that hold

  * ``compute_scattering_cffi`` (Beam.py:663-816, the kinematic sum), and
  * ``prop_mul_cpu`` (Beam.py:4850-4891, the propagation multiplier),

and pipes each straight into gcc (``-x c -``) with the reference's own
``-O3`` plus ``-march=x86-64-v3`` (the reference uses ``-march=native``,
hardware.py:108-124; the built library also runs on the host of the GPU
machine, whose CPU differs from this one, so a portable AVX2/FMA level is
used).  No source text is written anywhere; only the shared libraries land in
``oracle/_ref/`` (git-ignored).

Run:  python oracle/build_ref.py [/root/reference]
Exits quietly when the reference is absent (nothing to build).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_ref")

# (marker that opens the raw string, line range it must start in, library)
_KERNELS = [
    ('c_source = r"""', '"""', "compute_scattering_cffi", "libref_scatter.so"),
    ("source = r'''", "'''", "prop_mul_cpu", "libref_prop.so"),
]


def _extract(text, opener, closer, must_contain):
    start = 0
    while True:
        i = text.find(opener, start)
        if i < 0:
            raise RuntimeError(f"no string literal holding {must_contain} found")
        j = text.find(closer, i + len(opener))
        body = text[i + len(opener):j]
        if must_contain in body and "{" in body:
            return body
        start = j + len(closer)


def build(ref_root="/root/reference", quiet=False):
    beam_py = os.path.join(ref_root, "Beam.py")
    if not os.path.exists(beam_py):
        if not quiet:
            print("build_ref: reference not present; nothing to build")
        return False
    with open(beam_py) as fh:
        text = fh.read()
    os.makedirs(OUT, exist_ok=True)
    for opener, closer, name, lib in _KERNELS:
        src = _extract(text, opener, closer, name)
        out = os.path.join(OUT, lib)
        cmd = ["gcc", "-O3", "-march=x86-64-v3", "-shared", "-fPIC", "-x", "c", "-",
               "-o", out, "-lm"]
        subprocess.run(cmd, input=src.encode(), check=True)
        if not quiet:
            print(f"build_ref: {name} -> {out}")
    return True


if __name__ == "__main__":
    build(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
