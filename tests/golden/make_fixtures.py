"""Writes the byte fixtures under tests/golden/.

vector_n3.bin encodes the vector file format of the reference's
mpla_save_vector (src/mpla.cpp:638-640): a native-endian int32 element count
followed by the raw float64 values.  The values are chosen by hand; the file
is data, not code.  Run: python tests/golden/make_fixtures.py
"""
import os
import struct

HERE = os.path.dirname(os.path.abspath(__file__))

VECTOR_N3 = [1.0, -2.5, 3.25]


def main():
    with open(os.path.join(HERE, "vector_n3.bin"), "wb") as f:
        f.write(struct.pack("<i", len(VECTOR_N3)))
        f.write(struct.pack("<%dd" % len(VECTOR_N3), *VECTOR_N3))


if __name__ == "__main__":
    main()
