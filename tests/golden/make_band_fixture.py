"""Writes tests/golden/MoS2_head.band: the header line, the first 20 k-point
records and the last record of the reference's bands.x file W/MoS2.band
(nbnd = 60, one k line and six energy lines per record).  Data only: the
readers' test parses it with the drivers' Fortran reader and compares with a
free-format parse of the same text."""

import os

SRC = "/root/reference/W/MoS2.band"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "MoS2_head.band")
PER_K = 7  # k line + 60 energies in lines of 10

lines = open(SRC).read().splitlines()
head, body = lines[0], lines[1:]
recs = [body[i:i + PER_K] for i in range(0, len(body), PER_K)]
keep = recs[:20] + recs[-1:]
with open(DST, "w") as f:
    f.write(head + "\n")
    for r in keep:
        f.write("\n".join(r) + "\n")
print(DST, len(recs), "records in the source,", len(keep), "kept")
