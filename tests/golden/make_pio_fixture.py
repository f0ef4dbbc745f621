"""Writes tests/golden/pio_diffusion/state#000000: the first 256 records of
the reference's test/tests/diffusionTest/snapshot.000000000001/state#000000
(a FIXRECORDASCII pio state file, fields gid Vm, lrec 35, on a 40^3 grid)
with its header unchanged except nrecord, which becomes 256.  The file's Vm
of cell gid is (gid % nx + gid / nx % ny + gid / (nx ny)) % 3
(test/tests/diffusionTest/initialConditions.py), which the reader test
checks for every record.

    python tests/golden/make_pio_fixture.py
"""
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "/root/reference/test/tests/diffusionTest/snapshot.000000000001/state#000000"
NREC = 256


def main():
    raw = open(SRC, "rb").read()
    end = raw.index(b"}\n\n") + 3
    header = raw[:end].decode()
    lrec = int(re.search(r"lrec = (\d+);", header).group(1))
    header = re.sub(r"nrecord = \d+;", f"nrecord = {NREC};", header)
    body = raw[end:end + NREC * lrec]
    assert len(body) == NREC * lrec
    os.makedirs(os.path.join(HERE, "pio_diffusion"), exist_ok=True)
    with open(os.path.join(HERE, "pio_diffusion", "state#000000"), "wb") as fh:
        fh.write(header.encode() + body)


if __name__ == "__main__":
    main()
