"""Writes prepare_b_known_answer.json: the known-answer data of the reference's
native smoke test (native/cuda/tests/prepare_b_smoke.cpp) -- issuer key, 64-byte
header, four message scalars, the six generator points and scalars of the
equivalent MSM, the expected compressed B, and the scalar r used as the
non-canonical boundary case.  Reads the reference tree at /root/reference."""
import json
import os
import re

SRC = "/root/reference/native/cuda/tests/prepare_b_smoke.cpp"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "prepare_b_known_answer.json")


def main():
    src = open(SRC).read()

    def one(name):
        return re.search(name + r'\s*=\s*(?:hex_to_bytes\()?\s*"([0-9a-f]+)"', src).group(1)

    def block(name):
        return re.findall(r'"([0-9a-f]+)"', re.search(name + r" = \{(.*?)\};", src, re.S).group(1))

    d = {"source": "native/cuda/tests/prepare_b_smoke.cpp (reference known answers)",
         "public_key": one("public_key"), "header": one("header"),
         "message_scalars": block(r"std::vector<std::vector<uint8_t>> scalars"),
         "msm_points": block("msm_points"), "msm_scalars": block("msm_scalars"),
         "expected_b": re.search(r'expected_b =\s*"([0-9a-f]+)"', src).group(1),
         "invalid_fr": re.search(r'invalid_fr = hex_to_bytes\(\s*"([0-9a-f]+)"', src).group(1)}
    with open(OUT, "w") as fh:
        json.dump(d, fh, indent=1)


if __name__ == "__main__":
    main()
