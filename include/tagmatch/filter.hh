// 192-bit filter of the TagMatch API (reference filter.hh:7, bitvector.hh).
//
// Same bit layout and memory image as the reference's bitvector<192>: three
// little-endian 64-bit blocks, string position p is bit (p % 64) of block
// (p / 64), i.e. bit (p % 32) of 32-bit word (p / 32)
// (reference bitvector.hh:104-139, :210-212, :506-541).
#ifndef TAGMATCH_AMD_FILTER_HH_INCLUDED
#define TAGMATCH_AMD_FILTER_HH_INCLUDED

#include <cstdint>
#include <cstring>
#include <cctype>
#include <istream>
#include <ostream>
#include <string>

class filter192 {
public:
	static const unsigned int WIDTH = 192;
	static const unsigned int BLOCKS = 3;

	filter192() { clear(); }
	explicit filter192(const std::string & s) {
		clear();
		unsigned p = 0;
		for (char c : s) {
			if (p >= WIDTH) break;
			if (c == '1') set_bit(p);
			++p;
		}
	}
	explicit filter192(const uint64_t * blocks) { std::memcpy(b_, blocks, sizeof(b_)); }

	static filter192 from_words(const uint32_t * w) {
		filter192 f;
		for (unsigned i = 0; i < BLOCKS; ++i) f.b_[i] = (uint64_t)w[2 * i] | ((uint64_t)w[2 * i + 1] << 32);
		return f;
	}

	void clear() noexcept { b_[0] = b_[1] = b_[2] = 0; }
	void fill() noexcept { b_[0] = b_[1] = b_[2] = ~0ULL; }
	bool is_empty() const noexcept { return !(b_[0] | b_[1] | b_[2]); }

	void set_bit(unsigned pos) noexcept { b_[pos / 64] |= 1ULL << (pos % 64); }
	bool operator[](unsigned pos) const noexcept { return (b_[pos / 64] >> (pos % 64)) & 1; }

	const uint64_t * begin64() const noexcept { return b_; }
	const uint64_t * end64() const noexcept { return b_ + BLOCKS; }

	uint32_t uint32_value(unsigned i) const noexcept {
		return (i % 2) ? (uint32_t)(b_[i / 2] >> 32) : (uint32_t)(b_[i / 2] & 0xffffffffULL);
	}
	void copy_into_uint32_array(uint32_t * x) const noexcept {
		for (unsigned i = 0; i < 6; ++i) x[i] = uint32_value(i);
	}

	filter192 & operator|=(const filter192 & x) noexcept { for (unsigned i = 0; i < BLOCKS; ++i) b_[i] |= x.b_[i]; return *this; }
	filter192 & operator&=(const filter192 & x) noexcept { for (unsigned i = 0; i < BLOCKS; ++i) b_[i] &= x.b_[i]; return *this; }
	filter192 & operator^=(const filter192 & x) noexcept { for (unsigned i = 0; i < BLOCKS; ++i) b_[i] ^= x.b_[i]; return *this; }
	bool operator==(const filter192 & x) const noexcept { return b_[0] == x.b_[0] && b_[1] == x.b_[1] && b_[2] == x.b_[2]; }
	bool operator!=(const filter192 & x) const noexcept { return !(*this == x); }

	// *this is a subset of x (reference bitvector.hh:290-296)
	bool subset_of(const filter192 & x) const noexcept {
		return !((b_[0] & ~x.b_[0]) | (b_[1] & ~x.b_[1]) | (b_[2] & ~x.b_[2]));
	}

	// lexicographic order of the bit strings, position 0 first
	// (reference bitvector.hh:389-407)
	bool operator<(const filter192 & x) const noexcept {
		for (unsigned i = 0; i < BLOCKS; ++i)
			if (b_[i] != x.b_[i]) return (x.b_[i] >> __builtin_ctzll(b_[i] ^ x.b_[i])) & 1;
		return false;
	}
	bool operator>(const filter192 & x) const noexcept { return x < *this; }

	// first differing position, WIDTH if equal (reference bitvector.hh:412-418)
	unsigned leftmost_diff(const filter192 & x) const noexcept {
		for (unsigned i = 0; i < BLOCKS; ++i)
			if (b_[i] != x.b_[i]) return i * 64 + __builtin_ctzll(b_[i] ^ x.b_[i]);
		return WIDTH;
	}

	// first set position >= pos, WIDTH if none (reference bitvector.hh:440-456)
	unsigned next_bit(unsigned pos) const noexcept {
		for (unsigned i = pos / 64; i < BLOCKS; ++i) {
			uint64_t v = b_[i];
			if (i == pos / 64) v &= ~0ULL << (pos % 64);
			if (v) return i * 64 + __builtin_ctzll(v);
		}
		return WIDTH;
	}

	unsigned popcount() const noexcept {
		return __builtin_popcountll(b_[0]) + __builtin_popcountll(b_[1]) + __builtin_popcountll(b_[2]);
	}

	// 24-byte little-endian image (reference bitvector.hh:506-541)
	std::ostream & write_binary(std::ostream & out) const {
		unsigned char buf[24];
		for (unsigned i = 0; i < 24; ++i) buf[i] = (unsigned char)(b_[i / 8] >> (8 * (i % 8)));
		return out.write(reinterpret_cast<const char *>(buf), 24);
	}
	std::istream & read_binary(std::istream & in) {
		unsigned char buf[24];
		if (in.read(reinterpret_cast<char *>(buf), 24)) {
			clear();
			for (unsigned i = 0; i < 24; ++i) b_[i / 8] |= (uint64_t)buf[i] << (8 * (i % 8));
		}
		return in;
	}
	// '0'/'1' characters, position 0 first (reference bitvector.hh:543-580)
	std::ostream & write_ascii(std::ostream & out) const {
		for (unsigned p = 0; p < WIDTH; ++p) out.put((*this)[p] ? '1' : '0');
		return out;
	}
	std::istream & read_ascii(std::istream & in) {
		while (std::isspace(in.peek())) in.get();
		clear();
		for (unsigned p = 0; p < WIDTH; ++p) {
			int c = in.peek();
			if (c != '0' && c != '1') break;
			in.get();
			if (c == '1') set_bit(p);
		}
		return in;
	}

private:
	uint64_t b_[3];
};

typedef filter192 filter_t;
typedef uint8_t filter_pos_t;

#endif
