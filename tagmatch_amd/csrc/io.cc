// Reference file formats for filter tables and queries, so that datasets
// written by the reference's tools can be fed to this engine and back.
//
//   filter  binary : 24 bytes, the six 32-bit words little-endian (the
//                    reference's bitvector<192> image, bitvector.hh:506-540);
//           ASCII  : 192 characters '0'/'1', bit 0 first (bitvector.hh:543-580);
//                    reading stops at the first other character and leaves the
//                    remaining bits zero.
//   FIB entry      : binary filter, u32 key count, u32 keys (fib.cc:14-35);
//                    ASCII "+ <filter> key key ..." (fib.cc:38-46).
//   partitioned    : u32 partition id, then a FIB entry (fib.hh:41-56);
//                    ASCII "f <partition> <filter> key key ..." (fib.cc:100-131).
//   prefix         : binary filter, u8 length, u32 partition, u32 size
//                    (fib.hh:58-81); ASCII "p <partition> <first length bits>
//                    <size>" (fib.cc:72-98).
//   query          : binary filter (query.cc:11-19); ASCII "! <filter>" (query.cc:21-40).
//
// Integers are little-endian, as the reference writes them on little-endian
// hosts (io_util.hh).  Host code only: no device is needed.
#include "../../include/tagmatch_amd.h"
#include "../../include/tagmatch/filter.hh"

#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace tmx {
int set_error(int code, const char * fmt, ...);   // hip_util.hip: text for tm_last_error()
}

#define TM_CHECK_IO(x)              \
	do {                            \
		const int rc_ = (x);        \
		if (rc_ != TM_OK) return rc_; \
	} while (0)

namespace {

struct file {
	std::FILE * f = nullptr;
	file(const char * path, const char * mode) : f(path ? std::fopen(path, mode) : nullptr) {}
	~file() {
		if (f) std::fclose(f);
	}
};

bool get_u32(std::FILE * f, uint32_t & v) {
	unsigned char b[4];
	if (std::fread(b, 1, 4, f) != 4) return false;
	v = (uint32_t)b[0] | ((uint32_t)b[1] << 8) | ((uint32_t)b[2] << 16) | ((uint32_t)b[3] << 24);
	return true;
}

bool put_u32(std::FILE * f, uint32_t v) {
	const unsigned char b[4] = {(unsigned char)v, (unsigned char)(v >> 8), (unsigned char)(v >> 16),
								(unsigned char)(v >> 24)};
	return std::fwrite(b, 1, 4, f) == 4;
}

bool get_filter(std::FILE * f, uint32_t * w) {
	for (int j = 0; j < 6; ++j)
		if (!get_u32(f, w[j])) return false;
	return true;
}

bool put_filter(std::FILE * f, const uint32_t * w) {
	for (int j = 0; j < 6; ++j)
		if (!put_u32(f, w[j])) return false;
	return true;
}

// ASCII filter at s[pos...]: whitespace, then '0'/'1' up to 192 bits; pos
// ends after the last bit read (or after the character that ended them).
void parse_bits(const std::string & s, size_t & pos, uint32_t * w) {
	for (int j = 0; j < 6; ++j) w[j] = 0;
	while (pos < s.size() && std::isspace((unsigned char)s[pos])) ++pos;
	for (int i = 0; i < 192; ++i) {
		if (pos >= s.size()) return;
		const char c = s[pos++];
		if (c == '1') w[i / 32] |= 1u << (i % 32);
		else if (c != '0') return;
	}
}

std::string bits_text(const uint32_t * w) {
	std::string s(192, '0');
	for (int i = 0; i < 192; ++i)
		if ((w[i / 32] >> (i % 32)) & 1u) s[i] = '1';
	return s;
}

bool read_line(std::FILE * f, std::string & line) {
	line.clear();
	int c;
	while ((c = std::fgetc(f)) != EOF) {
		if (c == '\n') return true;
		line.push_back((char)c);
	}
	return !line.empty();
}

// A line's command ("+", "!", ...) and the position after it.
std::string command_of(const std::string & line, size_t & pos) {
	pos = 0;
	while (pos < line.size() && std::isspace((unsigned char)line[pos])) ++pos;
	const size_t b = pos;
	while (pos < line.size() && !std::isspace((unsigned char)line[pos])) ++pos;
	return line.substr(b, pos - b);
}

// Keys after the filter of an ASCII FIB line.
void parse_keys(const std::string & line, size_t pos, std::vector<uint32_t> & keys) {
	keys.clear();
	const char * p = line.c_str() + pos;
	char * end = nullptr;
	for (;;) {
		while (*p && std::isspace((unsigned char)*p)) ++p;
		if (!*p) return;
		const unsigned long v = std::strtoul(p, &end, 10);
		if (end == p) return;   // not a number: the entry ends
		keys.push_back((uint32_t)v);
		p = end;
	}
}

int check_format(uint32_t format, bool fib) {
	if (format == TM_IO_BINARY || format == TM_IO_ASCII) return TM_OK;
	if (fib && (format == TM_IO_BINARY_PARTITIONED || format == TM_IO_ASCII_PARTITIONED)) return TM_OK;
	return tmx::set_error(TM_ERR_INVALID, "unknown file format %u", format);
}

bool is_text(uint32_t format) { return format == TM_IO_ASCII || format == TM_IO_ASCII_PARTITIONED; }

// A decimal integer at line[pos...] (after whitespace); false if none.
bool parse_uint(const std::string & line, size_t & pos, uint32_t & v) {
	while (pos < line.size() && std::isspace((unsigned char)line[pos])) ++pos;
	const char * p = line.c_str() + pos;
	char * end = nullptr;
	const unsigned long x = std::strtoul(p, &end, 10);
	if (end == p) return false;
	v = (uint32_t)x;
	pos += (size_t)(end - p);
	return true;
}

// One FIB entry in `format`.
bool put_fib_entry(std::FILE * f, uint32_t format, const uint32_t * w, const uint32_t * keys, uint32_t n,
				   uint32_t part) {
	bool ok = true;
	if (is_text(format)) {
		ok = format == TM_IO_ASCII ? std::fprintf(f, "+ %s", bits_text(w).c_str()) > 0
								   : std::fprintf(f, "f %u %s", part, bits_text(w).c_str()) > 0;
		for (uint32_t x = 0; ok && x < n; ++x) ok = std::fprintf(f, " %u", keys[x]) > 0;
		return ok && std::fputc('\n', f) != EOF;
	}
	if (format == TM_IO_BINARY_PARTITIONED) ok = put_u32(f, part);
	ok = ok && put_filter(f, w) && put_u32(f, n);
	for (uint32_t x = 0; ok && x < n; ++x) ok = put_u32(f, keys[x]);
	return ok;
}

bool put_query(std::FILE * f, uint32_t format, const uint32_t * w) {
	return format == TM_IO_ASCII ? std::fprintf(f, "! %s\n", bits_text(w).c_str()) > 0 : put_filter(f, w);
}

bool put_prefix(std::FILE * f, uint32_t format, const tm_prefix & p) {
	if (format == TM_IO_ASCII) {
		std::string bits = bits_text(p.words);
		bits.resize(std::min<size_t>(p.length & 0xFF, 192));
		return std::fprintf(f, "p %u %s %u\n", p.partition, bits.c_str(), p.size) > 0;
	}
	const unsigned char len = (unsigned char)(p.length & 0xFF);
	return put_filter(f, p.words) && std::fwrite(&len, 1, 1, f) == 1 && put_u32(f, p.partition) &&
		   put_u32(f, p.size);
}

// Visit every FIB entry of a file: fn(words, keys, partition).
template <typename F>
int each_fib_entry(const char * path, uint32_t format, F && fn) {
	TM_CHECK_IO(check_format(format, true));
	file in(path, is_text(format) ? "r" : "rb");
	if (!in.f) return tmx::set_error(TM_ERR_INVALID, "cannot open %s", path ? path : "(null)");
	uint32_t w[6];
	std::vector<uint32_t> keys;
	if (is_text(format)) {
		const char * want = format == TM_IO_ASCII ? "+" : "f";
		std::string line;
		while (read_line(in.f, line)) {
			size_t pos;
			if (command_of(line, pos) != want) continue;   // other commands: not entries
			uint32_t part = 0;
			if (format == TM_IO_ASCII_PARTITIONED && !parse_uint(line, pos, part)) continue;
			parse_bits(line, pos, w);
			parse_keys(line, pos, keys);
			fn(w, keys, part);
		}
		return TM_OK;
	}
	uint64_t at = 0;
	for (;;) {
		uint32_t part = 0;
		if (format == TM_IO_BINARY_PARTITIONED) {
			if (!get_u32(in.f, part)) break;   // clean end of file
			if (!get_filter(in.f, w))
				return tmx::set_error(TM_ERR_INVALID, "%s: entry %llu truncated", path, (unsigned long long)at);
		} else if (!get_filter(in.f, w)) {
			break;
		}
		uint32_t n = 0;
		if (!get_u32(in.f, n))
			return tmx::set_error(TM_ERR_INVALID, "%s: entry %llu truncated", path, (unsigned long long)at);
		keys.resize(n);
		for (uint32_t i = 0; i < n; ++i)
			if (!get_u32(in.f, keys[i]))
				return tmx::set_error(TM_ERR_INVALID, "%s: entry %llu truncated", path, (unsigned long long)at);
		fn(w, keys, part);
		++at;
	}
	return TM_OK;
}

} // namespace

extern "C" {

int tm_io_scan_fib(const char * path, uint32_t format, uint64_t * n_filters, uint64_t * n_keys) {
	if (!n_filters || !n_keys) return tmx::set_error(TM_ERR_INVALID, "NULL argument");
	uint64_t nf = 0, nk = 0;
	const int rc = each_fib_entry(path, format, [&](const uint32_t *, const std::vector<uint32_t> & k, uint32_t) {
		++nf;
		nk += k.size();
	});
	if (rc != TM_OK) return rc;
	*n_filters = nf;
	*n_keys = nk;
	return TM_OK;
}

int tm_io_read_fib(const char * path, uint32_t format, uint32_t * words, uint32_t * key_offsets, uint32_t * keys,
				   uint32_t * partitions, uint64_t n_filters, uint64_t n_keys) {
	if ((n_filters && !words) || !key_offsets || (n_keys && !keys)) return tmx::set_error(TM_ERR_INVALID, "NULL argument");
	if (n_keys > 0xFFFFFFFFull) return tmx::set_error(TM_ERR_INVALID, "more than 2^32 - 1 keys");
	uint64_t i = 0, k = 0;
	bool over = false;
	key_offsets[0] = 0;
	const int rc = each_fib_entry(path, format, [&](const uint32_t * w, const std::vector<uint32_t> & kv, uint32_t part) {
		if (i >= n_filters || k + kv.size() > n_keys) {
			over = true;
			return;
		}
		std::memcpy(words + i * 6, w, 24);
		if (!kv.empty()) std::memcpy(keys + k, kv.data(), kv.size() * 4);
		k += kv.size();
		if (partitions) partitions[i] = part;
		key_offsets[++i] = (uint32_t)k;
	});
	if (rc != TM_OK) return rc;
	if (over || i != n_filters || k != n_keys)
		return tmx::set_error(TM_ERR_INVALID, "%s holds another number of entries or keys than given (scan it first)",
							  path);
	return TM_OK;
}

int tm_io_write_fib(const char * path, uint32_t format, const uint32_t * words, uint64_t n,
					const uint32_t * key_offsets, const uint32_t * keys, const uint32_t * partitions) {
	TM_CHECK_IO(check_format(format, true));
	if (n && (!words || !key_offsets || (key_offsets[n] && !keys))) return tmx::set_error(TM_ERR_INVALID, "NULL argument");
	file out(path, is_text(format) ? "w" : "wb");
	if (!out.f) return tmx::set_error(TM_ERR_INVALID, "cannot create %s", path ? path : "(null)");
	for (uint64_t i = 0; i < n; ++i) {
		const uint32_t kb = key_offsets[i], ke = key_offsets[i + 1];
		if (ke < kb) return tmx::set_error(TM_ERR_INVALID, "key offsets decrease at entry %llu", (unsigned long long)i);
		if (!put_fib_entry(out.f, format, words + i * 6, keys + kb, ke - kb, partitions ? partitions[i] : 0u))
			return tmx::set_error(TM_ERR_INVALID, "%s: write failed at entry %llu", path, (unsigned long long)i);
	}
	return TM_OK;
}

int tm_io_scan_queries(const char * path, uint32_t format, uint64_t * n) {
	if (!n) return tmx::set_error(TM_ERR_INVALID, "NULL argument");
	TM_CHECK_IO(check_format(format, false));
	file in(path, format == TM_IO_ASCII ? "r" : "rb");
	if (!in.f) return tmx::set_error(TM_ERR_INVALID, "cannot open %s", path ? path : "(null)");
	uint64_t c = 0;
	if (format == TM_IO_ASCII) {
		std::string line;
		while (read_line(in.f, line)) {
			size_t pos;
			c += command_of(line, pos) == "!";
		}
	} else {
		if (std::fseek(in.f, 0, SEEK_END) != 0) return tmx::set_error(TM_ERR_INVALID, "%s: cannot seek", path);
		const long size = std::ftell(in.f);
		if (size < 0 || size % 24) return tmx::set_error(TM_ERR_INVALID, "%s: size is not a multiple of 24 bytes", path);
		c = (uint64_t)size / 24;
	}
	*n = c;
	return TM_OK;
}

int tm_io_read_queries(const char * path, uint32_t format, uint32_t * words, uint64_t n) {
	if (n && !words) return tmx::set_error(TM_ERR_INVALID, "NULL argument");
	TM_CHECK_IO(check_format(format, false));
	file in(path, format == TM_IO_ASCII ? "r" : "rb");
	if (!in.f) return tmx::set_error(TM_ERR_INVALID, "cannot open %s", path ? path : "(null)");
	uint64_t i = 0;
	if (format == TM_IO_ASCII) {
		std::string line;
		while (i < n && read_line(in.f, line)) {
			size_t pos;
			if (command_of(line, pos) != "!") continue;
			parse_bits(line, pos, words + i * 6);
			++i;
		}
	} else {
		while (i < n && get_filter(in.f, words + i * 6)) ++i;
	}
	if (i != n) return tmx::set_error(TM_ERR_INVALID, "%s holds %llu queries, not %llu", path, (unsigned long long)i,
									  (unsigned long long)n);
	return TM_OK;
}

int tm_io_write_queries(const char * path, uint32_t format, const uint32_t * words, uint64_t n) {
	if (n && !words) return tmx::set_error(TM_ERR_INVALID, "NULL argument");
	TM_CHECK_IO(check_format(format, false));
	file out(path, format == TM_IO_ASCII ? "w" : "wb");
	if (!out.f) return tmx::set_error(TM_ERR_INVALID, "cannot create %s", path ? path : "(null)");
	for (uint64_t i = 0; i < n; ++i) {
		const bool ok = format == TM_IO_ASCII ? std::fprintf(out.f, "! %s\n", bits_text(words + i * 6).c_str()) > 0
											  : put_filter(out.f, words + i * 6);
		if (!ok) return tmx::set_error(TM_ERR_INVALID, "%s: write failed at query %llu", path, (unsigned long long)i);
	}
	return TM_OK;
}

} // extern "C"

// ------------------------------------------------------------------ tools

namespace {

// Visit every query of a file: fn(words).
template <typename F>
int each_query(const char * path, uint32_t format, F && fn) {
	TM_CHECK_IO(check_format(format, false));
	file in(path, format == TM_IO_ASCII ? "r" : "rb");
	if (!in.f) return tmx::set_error(TM_ERR_INVALID, "cannot open %s", path ? path : "(null)");
	uint32_t w[6];
	if (format == TM_IO_ASCII) {
		std::string line;
		while (read_line(in.f, line)) {
			size_t pos;
			if (command_of(line, pos) != "!") continue;
			parse_bits(line, pos, w);
			if (!fn(w)) break;
		}
	} else {
		while (get_filter(in.f, w))
			if (!fn(w)) break;
	}
	return TM_OK;
}

// Visit every partition prefix of a file: fn(prefix).
template <typename F>
int each_prefix(const char * path, uint32_t format, F && fn) {
	TM_CHECK_IO(check_format(format, false));
	file in(path, format == TM_IO_ASCII ? "r" : "rb");
	if (!in.f) return tmx::set_error(TM_ERR_INVALID, "cannot open %s", path ? path : "(null)");
	tm_prefix p;
	std::memset(&p, 0, sizeof(p));
	if (format == TM_IO_ASCII) {
		std::string line;
		while (read_line(in.f, line)) {
			size_t pos;
			if (command_of(line, pos) != "p" || !parse_uint(line, pos, p.partition)) continue;
			while (pos < line.size() && std::isspace((unsigned char)line[pos])) ++pos;
			size_t end = pos;
			while (end < line.size() && !std::isspace((unsigned char)line[end])) ++end;
			p.length = (uint32_t)(uint8_t)(end - pos);   // the bit string's length, as a uint8_t
			size_t bp = pos;
			parse_bits(line.substr(0, end), bp, p.words);
			pos = end;
			parse_uint(line, pos, p.size);
			fn(p);
		}
		return TM_OK;
	}
	uint64_t at = 0;
	for (;;) {
		if (!get_filter(in.f, p.words)) break;
		unsigned char len = 0;
		if (std::fread(&len, 1, 1, in.f) != 1 || !get_u32(in.f, p.partition) || !get_u32(in.f, p.size))
			return tmx::set_error(TM_ERR_INVALID, "%s: prefix %llu truncated", path, (unsigned long long)at);
		p.length = len;
		fn(p);
		++at;
	}
	return TM_OK;
}

// std::default_random_engine of libstdc++ (minstd_rand0: x <- 16807 x mod
// 2^31 - 1, seeded with s mod m, 0 replaced by 1) and the draws of its
// uniform_int_distribution over [0, n] for a generator of range
// [1, 2^31 - 2]: rejection of the draws at or above the largest multiple of
// (n + 1) that fits the generator's range, then division.
struct minstd0 {
	uint64_t x;
	explicit minstd0(uint32_t seed) : x(seed % 2147483647u) {
		if (x == 0) x = 1;
	}
	uint64_t next() { return x = x * 16807u % 2147483647u; }
	uint64_t uniform(uint64_t n) {   // [0, n], n < 2^31 - 2
		const uint64_t range = 2147483646u - 1u;
		const uint64_t span = n + 1;
		const uint64_t scaling = range / span;
		const uint64_t past = span * scaling;
		uint64_t r;
		do r = next() - 1u;
		while (r >= past);
		return r / scaling;
	}
};

// The reference tool's sampler: keep a record with probability (wanted) /
// (left), drawn as uniform(0, left - 1) < wanted (sample_workload.cc:24-56).
struct selection {
	minstd0 g;
	uint64_t n, k;
	selection(uint32_t seed, uint64_t n_, uint64_t k_) : g(seed), n(n_), k(k_) {}
	bool take() {
		if (k == 0) return false;
		if (k < n) {
			--n;
			if (g.uniform(n) < k) {
				--k;
				return true;
			}
			return false;
		}
		--k;
		return true;
	}
};

struct fib_rec {
	filter192 f;
	uint32_t w[6];
	std::vector<uint32_t> keys;
};

} // namespace

extern "C" {

int tm_io_scan_prefixes(const char * path, uint32_t format, uint64_t * n) {
	if (!n) return tmx::set_error(TM_ERR_INVALID, "NULL argument");
	uint64_t c = 0;
	TM_CHECK_IO(each_prefix(path, format, [&](const tm_prefix &) { ++c; }));
	*n = c;
	return TM_OK;
}

int tm_io_read_prefixes(const char * path, uint32_t format, tm_prefix * prefixes, uint64_t n) {
	if (n && !prefixes) return tmx::set_error(TM_ERR_INVALID, "NULL argument");
	uint64_t i = 0;
	TM_CHECK_IO(each_prefix(path, format, [&](const tm_prefix & p) {
		if (i < n) prefixes[i] = p;
		++i;
	}));
	if (i != n) return tmx::set_error(TM_ERR_INVALID, "%s holds %llu prefixes, not %llu", path, (unsigned long long)i,
									  (unsigned long long)n);
	return TM_OK;
}

int tm_io_write_prefixes(const char * path, uint32_t format, const tm_prefix * prefixes, uint64_t n) {
	if (n && !prefixes) return tmx::set_error(TM_ERR_INVALID, "NULL argument");
	TM_CHECK_IO(check_format(format, false));
	file out(path, format == TM_IO_ASCII ? "w" : "wb");
	if (!out.f) return tmx::set_error(TM_ERR_INVALID, "cannot create %s", path ? path : "(null)");
	for (uint64_t i = 0; i < n; ++i)
		if (!put_prefix(out.f, format, prefixes[i]))
			return tmx::set_error(TM_ERR_INVALID, "%s: write failed at prefix %llu", path, (unsigned long long)i);
	return TM_OK;
}

int tm_io_convert(const char * in, uint32_t in_format, const char * out, uint32_t out_format, uint32_t kind,
				  uint64_t * n) {
	const bool fib = kind == TM_IO_KIND_FIB;
	if (kind > TM_IO_KIND_PREFIXES) return tmx::set_error(TM_ERR_INVALID, "unknown record kind %u", kind);
	TM_CHECK_IO(check_format(in_format, fib));
	TM_CHECK_IO(check_format(out_format, fib));
	file o(out, is_text(out_format) ? "w" : "wb");
	if (!o.f) return tmx::set_error(TM_ERR_INVALID, "cannot create %s", out ? out : "(null)");
	uint64_t c = 0;
	bool ok = true;
	int rc;
	if (fib) {
		rc = each_fib_entry(in, in_format, [&](const uint32_t * w, const std::vector<uint32_t> & k, uint32_t part) {
			ok = ok && put_fib_entry(o.f, out_format, w, k.data(), (uint32_t)k.size(), part);
			++c;
		});
	} else if (kind == TM_IO_KIND_QUERIES) {
		rc = each_query(in, in_format, [&](const uint32_t * w) {
			ok = ok && put_query(o.f, out_format, w);
			++c;
			return true;
		});
	} else {
		rc = each_prefix(in, in_format, [&](const tm_prefix & p) {
			ok = ok && put_prefix(o.f, out_format, p);
			++c;
		});
	}
	TM_CHECK_IO(rc);
	if (!ok) return tmx::set_error(TM_ERR_INVALID, "%s: write failed", out);
	if (n) *n = c;
	return TM_OK;
}

int tm_io_sample(const char * in, uint32_t format, const char * out, uint32_t kind, uint64_t n_total, uint64_t k,
				 uint32_t seed, uint64_t * n_out) {
	if (kind != TM_IO_KIND_FIB && kind != TM_IO_KIND_QUERIES)
		return tmx::set_error(TM_ERR_INVALID, "sampling takes FIB entries or queries");
	if (n_total == 0 || k == 0) return tmx::set_error(TM_ERR_INVALID, "n_total and k must be > 0");
	if (n_total >= 2147483646u) return tmx::set_error(TM_ERR_INVALID, "n_total must be below 2^31 - 2");
	TM_CHECK_IO(check_format(format, kind == TM_IO_KIND_FIB));
	file o(out, is_text(format) ? "w" : "wb");
	if (!o.f) return tmx::set_error(TM_ERR_INVALID, "cannot create %s", out ? out : "(null)");
	selection sel(seed, n_total, k);
	uint64_t c = 0;
	bool ok = true;
	int rc;
	if (kind == TM_IO_KIND_FIB) {
		// stops reading once k entries are kept, as the reference's loop does
		bool done = false;
		rc = each_fib_entry(in, format, [&](const uint32_t * w, const std::vector<uint32_t> & keys, uint32_t part) {
			if (done || sel.k == 0) {
				done = true;
				return;
			}
			if (sel.take()) {
				ok = ok && put_fib_entry(o.f, format, w, keys.data(), (uint32_t)keys.size(), part);
				++c;
			}
		});
	} else {
		rc = each_query(in, format, [&](const uint32_t * w) {
			if (sel.take()) {
				ok = ok && put_query(o.f, format, w);
				++c;
			}
			return true;
		});
	}
	TM_CHECK_IO(rc);
	if (!ok) return tmx::set_error(TM_ERR_INVALID, "%s: write failed", out);
	if (n_out) *n_out = c;
	return TM_OK;
}

int tm_io_split(const char * in, uint32_t format, const char * prefixes_out, const char * filters_out,
				uint32_t max_size, uint32_t min_k, uint64_t * n_parts) {
	if (format != TM_IO_BINARY && format != TM_IO_ASCII)
		return tmx::set_error(TM_ERR_INVALID, "split reads TM_IO_BINARY or TM_IO_ASCII entries");
	std::vector<fib_rec> fib;
	bool ascending = true;
	TM_CHECK_IO(each_fib_entry(in, format, [&](const uint32_t * w, const std::vector<uint32_t> & keys, uint32_t) {
		fib_rec r;
		std::memcpy(r.w, w, 24);
		r.f = filter192::from_words(w);
		r.keys = keys;
		if (ascending && !fib.empty()) ascending = fib.back().f < r.f;
		fib.push_back(std::move(r));
	}));
	// the reference keeps a strictly ascending input as it is and otherwise
	// sorts it in descending order; equal filters keep their input order
	if (!ascending)
		std::stable_sort(fib.begin(), fib.end(), [](const fib_rec & a, const fib_rec & b) { return b.f < a.f; });
	file po(prefixes_out, format == TM_IO_ASCII ? "w" : "wb");
	file fo(filters_out, format == TM_IO_ASCII ? "w" : "wb");
	if (prefixes_out && !po.f) return tmx::set_error(TM_ERR_INVALID, "cannot create %s", prefixes_out);
	if (filters_out && !fo.f) return tmx::set_error(TM_ERR_INVALID, "cannot create %s", filters_out);
	const uint32_t fmt_f = format == TM_IO_ASCII ? TM_IO_ASCII_PARTITIONED : TM_IO_BINARY_PARTITIONED;
	const size_t n = fib.size();
	uint32_t pid = 0;
	size_t f = 0;
	while (f < n) {
		tm_prefix p;
		std::memcpy(p.words, fib[f].w, 24);
		p.partition = pid;
		uint8_t length = 192;
		// position of the (min_k + 1)-th set bit of the first filter (192 if none)
		unsigned kth = fib[f].f.next_bit(0);
		for (uint32_t i = 0; i < min_k && kth < 192; ++i) kth = fib[f].f.next_bit(kth + 1);
		size_t g = f + 1, next_f = g;
		while (g < n && g - f < max_size) {
			const unsigned msd = fib[f].f.leftmost_diff(fib[g].f);
			if (msd < length) {
				if (kth >= msd) {
					next_f = g;
					++length;
					break;
				}
				length = (uint8_t)msd;
				next_f = g;
			}
			++g;
		}
		if (g == n) {
			next_f = g;
			--length;
		}
		p.length = length;
		p.size = (uint32_t)(next_f - f);
		if (po.f && !put_prefix(po.f, format, p))
			return tmx::set_error(TM_ERR_INVALID, "%s: write failed", prefixes_out);
		for (; f < next_f; ++f)
			if (fo.f && !put_fib_entry(fo.f, fmt_f, fib[f].w, fib[f].keys.data(), (uint32_t)fib[f].keys.size(), pid))
				return tmx::set_error(TM_ERR_INVALID, "%s: write failed", filters_out);
		++pid;
	}
	if (n_parts) *n_parts = pid;
	return TM_OK;
}

} // extern "C"
