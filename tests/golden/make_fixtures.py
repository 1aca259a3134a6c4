"""Writes the reference's console_test MRC fixtures as data files under tests/golden/.

The reference embeds three MRC files as C byte arrays in
src/programs/console_test/{hiv_image_80x80x1,hiv_images_shift_noise_80x80x10,sine_128x128x1}.cpp
and writes them to disk before its known-answer tests run (console_test.cpp:1086-1126).
This script parses those arrays (text) and writes the bytes unchanged as .mrc files, so the
tests can use the same inputs without the reference tree. Run from the repository root:

    python tests/golden/make_fixtures.py /root/reference
"""
import os
import re
import sys

FIXTURES = {
    "hiv_image_80x80x1.cpp": "hiv_image_80x80x1.mrc",
    "hiv_images_shift_noise_80x80x10.cpp": "hiv_images_shift_noise_80x80x10.mrc",
    "sine_128x128x1.cpp": "sine_128x128x1.mrc",
}


def main(reference_root: str) -> None:
    src_dir = os.path.join(reference_root, "src", "programs", "console_test")
    out_dir = os.path.dirname(os.path.abspath(__file__))
    for src, dst in FIXTURES.items():
        text = open(os.path.join(src_dir, src)).read()
        body = text[text.index("{") + 1: text.rindex("}")]
        data = bytes(int(tok, 16) for tok in re.findall(r"0x[0-9a-fA-F]{2}", body))
        with open(os.path.join(out_dir, dst), "wb") as f:
            f.write(data)
        print(f"{dst}: {len(data)} bytes")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
