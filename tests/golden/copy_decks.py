"""Copy the reference's channel input decks as fixtures.

The decks (examples/<case>/singleideal.ini) are the inputs of the reference's
own known-answer runs (tests/tests.ini patches them); the deck reader of this
library is checked against them.  Run once in a checkout that has the
reference: python tests/golden/copy_decks.py /path/to/reference
"""
import os
import shutil
import sys

CASES = ("subsonic_channel", "supersonic_channel")


def main(ref):
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "decks")
    os.makedirs(here, exist_ok=True)
    for case in CASES:
        shutil.copyfile(os.path.join(ref, "examples", case, "singleideal.ini"), os.path.join(here, f"{case}.ini"))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
