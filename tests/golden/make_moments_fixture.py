"""Write moments_im_2pop.json from the reference's cached moments output.

The reference caches moments' two-population LD parse of
``examples/data/im-parsing-example.vcf`` (pops deme0/deme1,
``bp_bins = logspace(2, 5, 4)``) in
``tests/cache/moments_ld_test_im_4bins_deme0-deme1.pkl``
(written by tests/test_ld_validation_full.py:38-72).

The pickle is only *disassembled* here with ``pickletools.genops``: its byte
payloads (little-endian float64 arrays) are decoded with ``struct``.  Nothing
in the file is executed and no object is reconstructed.

Run from the repository root in a container that has the reference mounted::

    python tests/golden/make_moments_fixture.py /root/reference
"""
import gzip
import json
import pickletools
import shutil
import struct
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent


def decode(pkl_bytes):
    strings, blobs = [], []
    for op, arg, _ in pickletools.genops(pkl_bytes):
        if op.name in ("SHORT_BINUNICODE", "BINUNICODE"):
            strings.append(arg)
        elif op.name in ("SHORT_BINBYTES", "BINBYTES"):
            blobs.append(arg)
    edges = [struct.unpack("<d", b)[0] for b in blobs if len(b) == 8]
    arrays = [list(struct.unpack("<%dd" % (len(b) // 8), b)) for b in blobs if len(b) > 8]
    ld = [s for s in strings if s.startswith(("DD_", "Dz_", "pi2_"))]
    het = [s for s in strings if s.startswith("H_")]
    pops = [s for s in strings if s.startswith("deme")]
    assert len(edges) == 6 and len(arrays) == 4, (len(edges), len(arrays))
    bins = [[edges[2 * k], edges[2 * k + 1]] for k in range(3)]
    return {
        "source": "tests/cache/moments_ld_test_im_4bins_deme0-deme1.pkl (moments output)",
        "vcf": "im-parsing-example.vcf.gz",
        "pop_file": "im_pop.txt",
        "pops": pops,
        "bp_bins": [100.0, 1000.0, 10000.0, 100000.0],
        "bins": bins,
        "ld_names": ld,
        "het_names": het,
        "sums": arrays[:3],
        "het": arrays[3],
    }


def main(ref_root):
    ref = Path(ref_root)
    out = decode((ref / "tests/cache/moments_ld_test_im_4bins_deme0-deme1.pkl").read_bytes())
    (HERE / "moments_im_2pop.json").write_text(json.dumps(out, indent=1))
    # The input data the fixture was computed on (reference test data, kept verbatim).
    with open(ref / "examples/data/im-parsing-example.vcf", "rb") as src, \
            gzip.GzipFile(HERE / "im-parsing-example.vcf.gz", "wb", mtime=0) as dst:
        shutil.copyfileobj(src, dst)
    shutil.copyfile(ref / "examples/data/im_pop.txt", HERE / "im_pop.txt")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
