#!/usr/bin/env python3
"""Regenerate the golden fixtures in this directory from the reference tree.

Run in a container where the reference checkout is mounted read-only at
/root/reference (it is not present on GPU hosts; the fixtures are committed).
Only data is taken: the reference's own test input and data files, and the
numbers its saved logs hold.  No reference source is copied.

Outputs
  batch0.fast5            test/batch0.fast5 (one multi-read FAST5 read)
  dna_kmer_model.txt      src/ref/dna_kmer_model.txt (r9.4 6-mer model)
  reference.fasta         src/ref/reference.fasta (lambda phage J02459.1)
  batch0_signal.npy       raw int16 Raw/Signal of the read in batch0.fast5
                          (h5dump binary dump, for CPU tests that do not use
                          the product's FAST5 reader)
  sdtw_log_golden.npz     scripts/log as numbers: query[4096], the six
                          printed columns D[i][0,1,31,32,R-2,R-1] (4096 x 6),
                          and the final min (286.8895631)
  fp16_score_logs.npz     src/mapped.log / src/unmapped.log FP16 kernel scores
                          (fwd, rev) as printed, for distribution-level checks
"""
import os
import re
import shutil
import subprocess
import sys

import numpy as np

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
READ = "read_0083a4a2-b062-44ed-988e-e7f5a301121f"


def main():
    if not os.path.isdir(REF):
        sys.exit("reference tree not mounted; fixtures are already committed")
    for src, dst in [("test/batch0.fast5", "batch0.fast5"),
                     ("src/ref/dna_kmer_model.txt", "dna_kmer_model.txt"),
                     ("src/ref/reference.fasta", "reference.fasta")]:
        shutil.copyfile(os.path.join(REF, src), os.path.join(HERE, dst))

    tmp = os.path.join("/tmp", "batch0_signal.bin")
    subprocess.run(["/opt/conda/bin/h5dump", "-d", f"/{READ}/Raw/Signal", "-b", "LE",
                    "-o", tmp, os.path.join(REF, "test/batch0.fast5")],
                   check=True, stdout=subprocess.DEVNULL)
    sig = np.fromfile(tmp, dtype="<i2")
    np.save(os.path.join(HERE, "batch0_signal.npy"), sig)

    q, cols, final = [], [], None
    pat = re.compile(r"i= (\d+) query= (\S+) val\[0\]= (\S+) val\[1\]= (\S+) "
                     r"val\[31\]= (\S+) val\[32\]= (\S+) val\[R-2\]= (\S+) val\[R-1\]= (\S+)")
    with open(os.path.join(REF, "scripts/log")) as f:
        for line in f:
            m = pat.match(line)
            if m:
                assert int(m.group(1)) == len(q)
                q.append(float(m.group(2)))
                cols.append([float(m.group(k)) for k in range(3, 9)])
            elif line.startswith("min scores"):
                final = float(re.search(r"array\(([^)]+)\)", line).group(1))
    np.savez_compressed(os.path.join(HERE, "sdtw_log_golden.npz"),
                        query=np.array(q, np.float64), cols=np.array(cols, np.float64),
                        col_index=np.array([0, 1, 31, 32, 96256 - 2, 96256 - 1], np.int64),
                        final_min=np.float64(final))

    def scores(name):
        out = []
        with open(os.path.join(REF, "src", name)) as f:
            for line in f:
                p = line.split()
                if len(p) == 6 and p[1].startswith("read_"):
                    out.append((float(p[4]), float(p[5])))
        return np.array(out, np.float64)

    np.savez_compressed(os.path.join(HERE, "fp16_score_logs.npz"),
                        mapped=scores("mapped.log"), unmapped=scores("unmapped.log"))
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
