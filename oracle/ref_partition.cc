// Driver for the reference partitioner (oracle/_ref only; test tooling).
//
// usage: ref_partition <filters.bin> <maxp> <threads> <out.bin>
//   filters.bin : n records of 24 bytes (the filter's binary image)
//   out.bin     : u32 n, then per final position i: u32 input index,
//                 u32 partition id; then u32 np, then per partition:
//                 24-byte mask, u32 id, u32 size
#include <cstdio>
#include <cstdint>
#include <cstdlib>
#include <fstream>
#include <vector>

#include "partitioning.hh"
#include "fib.hh"

int main(int argc, char ** argv) {
	if (argc != 5) {
		std::fprintf(stderr, "usage: %s filters.bin maxp threads out.bin\n", argv[0]);
		return 2;
	}
	std::ifstream in(argv[1], std::ios::binary);
	std::vector<partition_fib_entry *> fib;
	uint32_t idx = 0;
	for (;;) {
		fib_entry e;
		if (!e.filter.read_binary(in)) break;
		e.keys.push_back(idx++);   // key = input index, to trace positions
		fib.push_back(new partition_fib_entry(e));
	}
	std::vector<partition_prefix> masks;
	partitioning::set_maxp((unsigned)std::atoi(argv[2]));
	partitioning::set_cpu_threads((unsigned)std::atoi(argv[3]));
	partitioning::balanced_partitioning(fib, masks);
	std::ofstream out(argv[4], std::ios::binary);
	uint32_t n = (uint32_t)fib.size();
	out.write((const char *)&n, 4);
	for (partition_fib_entry * p : fib) {
		uint32_t k = p->keys[0], part = p->partition;
		out.write((const char *)&k, 4);
		out.write((const char *)&part, 4);
	}
	uint32_t np = (uint32_t)masks.size();
	out.write((const char *)&np, 4);
	for (const partition_prefix & m : masks) {
		m.filter.write_binary(out);
		out.write((const char *)&m.partition, 4);
		out.write((const char *)&m.size, 4);
	}
	for (partition_fib_entry * p : fib) delete p;
	return 0;
}
