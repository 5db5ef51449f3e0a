"""Write io_vectors.json: the filter encoding cases of the reference's
test/test_io.cc (an ASCII filter and the 24 bytes its binary form must
be), lifted as data.  Run in a checkout that has the reference at
/root/reference; the JSON is what the tests read."""
import json
import os
import re

SRC = "/root/reference/test/test_io.cc"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "io_vectors.json")


def c_bytes(lit):
    """The bytes of a C string literal body (octal escapes as the reference writes them)."""
    return lit.encode("latin1").decode("unicode_escape").encode("latin1")


def main():
    text = open(SRC).read()
    cases = []
    for name, ascii_, lit, size in re.findall(
            r'BOOST_AUTO_TEST_CASE\(\s*(\w+)\s*\)\s*\{\s*BOOST_CHECK\(check_binary_encoding\("([01]+)",\s*"([^"]*)",\s*(\d+)\)\);',
            text):
        b = c_bytes(lit)[:int(size)]
        cases.append({"name": name, "ascii": ascii_, "binary_hex": b.hex()})
    json.dump({"source": "reference test/test_io.cc (bitvector_io suite)", "cases": cases}, open(OUT, "w"), indent=1)
    print(f"{len(cases)} cases -> {OUT}")


if __name__ == "__main__":
    main()
