"""Copies (gzipped) the reference's tutorial models, pressed model files and
sequence files that the drop-in pipeline and pressed-file tests read: data
files of forestcheng/cudahmmsearch code/tutorial/, unchanged.

    python tests/golden/hmmer/make_inputs.py
"""
import gzip
import os
import shutil

TUT = "/root/reference/code/tutorial"
FILES = ["fn3.hmm", "globins4.hmm", "Pkinase.hmm", "MADE1.hmm",
         "7LESS_DROME", "HBB_HUMAN", "globins45.fa", "dna_target.fa",
         "minifam", "minifam.h3f", "minifam.h3p", "MADE1.hmm.h3f", "MADE1.hmm.h3p"]


def main():
    here = os.path.dirname(os.path.abspath(__file__))
    for f in FILES:
        with open(os.path.join(TUT, f), "rb") as src, gzip.GzipFile(os.path.join(here, f + ".gz"), "wb", mtime=0) as dst:
            shutil.copyfileobj(src, dst)
        print("wrote", f + ".gz")


if __name__ == "__main__":
    main()
