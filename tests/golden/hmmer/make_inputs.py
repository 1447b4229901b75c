"""Copies (gzipped) the reference's tutorial models, pressed model files and
sequence files that the drop-in pipeline, pressed-file and text-reader tests
read, and the text models of its testsuite: data files of
forestcheng/cudahmmsearch code/tutorial/ and code/testsuite/, unchanged.

    python tests/golden/hmmer/make_inputs.py
"""
import gzip
import os
import shutil

TUT = "/root/reference/code/tutorial"
TESTSUITE = "/root/reference/code/testsuite"
# HMMER3 text models of the reference's testsuite (copied as testsuite_<name>.gz)
TS_MODELS = ["20aa.hmm", "3box.hmm", "Caudal_act.hmm", "LuxC.hmm", "M1.hmm", "PSE.hmm", "Patched.hmm",
             "RRM_1.hmm", "SMC_N.hmm", "XYPPX.hmm", "ecori.hmm"]
FILES = ["fn3.hmm", "globins4.hmm", "Pkinase.hmm", "MADE1.hmm",
         "7LESS_DROME", "HBB_HUMAN", "globins45.fa", "dna_target.fa",
         "minifam", "minifam.h3f", "minifam.h3p", "MADE1.hmm.h3f", "MADE1.hmm.h3p"]


def main():
    here = os.path.dirname(os.path.abspath(__file__))
    for f in FILES:
        with open(os.path.join(TUT, f), "rb") as src, gzip.GzipFile(os.path.join(here, f + ".gz"), "wb", mtime=0) as dst:
            shutil.copyfileobj(src, dst)
        print("wrote", f + ".gz")
    for f in TS_MODELS:
        with open(os.path.join(TESTSUITE, f), "rb") as src, \
                gzip.GzipFile(os.path.join(here, "testsuite_" + f + ".gz"), "wb", mtime=0) as dst:
            shutil.copyfileobj(src, dst)
        print("wrote", "testsuite_" + f + ".gz")


if __name__ == "__main__":
    main()
