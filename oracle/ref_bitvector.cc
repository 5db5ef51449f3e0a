// Driver for the reference bitvector<192> (oracle/_ref only; test tooling).
//
// usage: ref_bitvector <pairs.bin> <out.bin>
//   pairs.bin : records of two 24-byte filters a, b
//   out.bin   : per record: u8 subset(a,b), u8 (a<b), u8 (b<a),
//               u8 leftmost_diff(a,b), u8 popcount(a), 3 pad bytes,
//               u32 next_bit(a, 0) and 6 x u32 uint32_value(a, j)
#include <cstdint>
#include <cstdio>
#include <fstream>

#include "bitvector.hh"

typedef bitvector<192> filter_t;

int main(int argc, char ** argv) {
	if (argc != 3) {
		std::fprintf(stderr, "usage: %s pairs.bin out.bin\n", argv[0]);
		return 2;
	}
	std::ifstream in(argv[1], std::ios::binary);
	std::ofstream out(argv[2], std::ios::binary);
	filter_t a, b;
	while (a.read_binary(in) && b.read_binary(in)) {
		uint8_t r[8] = {0};
		r[0] = a.subset_of(b);
		r[1] = a < b;
		r[2] = b < a;
		r[3] = (uint8_t)a.leftmost_diff(b);
		r[4] = (uint8_t)a.popcount();
		out.write((const char *)r, 8);
		uint32_t nb = a.next_bit(0);
		out.write((const char *)&nb, 4);
		for (unsigned j = 0; j < 6; ++j) {
			uint32_t w = a.uint32_value(j);
			out.write((const char *)&w, 4);
		}
	}
	return 0;
}
