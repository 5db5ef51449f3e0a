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
//   partitioned    : u32 partition id, then a FIB entry (fib.hh:41-56).
//   query          : binary filter (query.cc:11-19); ASCII "! <filter>" (query.cc:21-40).
//
// Integers are little-endian, as the reference writes them on little-endian
// hosts (io_util.hh).  Host code only: no device is needed.
#include "../../include/tagmatch_amd.h"

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
	if (fib && format == TM_IO_BINARY_PARTITIONED) return TM_OK;
	return tmx::set_error(TM_ERR_INVALID, "unknown file format %u", format);
}

// Visit every FIB entry of a file: fn(words, keys, partition).
template <typename F>
int each_fib_entry(const char * path, uint32_t format, F && fn) {
	TM_CHECK_IO(check_format(format, true));
	file in(path, format == TM_IO_ASCII ? "r" : "rb");
	if (!in.f) return tmx::set_error(TM_ERR_INVALID, "cannot open %s", path ? path : "(null)");
	uint32_t w[6];
	std::vector<uint32_t> keys;
	if (format == TM_IO_ASCII) {
		std::string line;
		while (read_line(in.f, line)) {
			size_t pos;
			if (command_of(line, pos) != "+") continue;   // other commands: not entries
			parse_bits(line, pos, w);
			parse_keys(line, pos, keys);
			fn(w, keys, 0u);
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
	file out(path, format == TM_IO_ASCII ? "w" : "wb");
	if (!out.f) return tmx::set_error(TM_ERR_INVALID, "cannot create %s", path ? path : "(null)");
	for (uint64_t i = 0; i < n; ++i) {
		const uint32_t * w = words + i * 6;
		const uint32_t kb = key_offsets[i], ke = key_offsets[i + 1];
		if (ke < kb) return tmx::set_error(TM_ERR_INVALID, "key offsets decrease at entry %llu", (unsigned long long)i);
		bool ok = true;
		if (format == TM_IO_ASCII) {
			ok = std::fprintf(out.f, "+ %s", bits_text(w).c_str()) > 0;
			for (uint32_t x = kb; ok && x < ke; ++x) ok = std::fprintf(out.f, " %u", keys[x]) > 0;
			ok = ok && std::fputc('\n', out.f) != EOF;
		} else {
			if (format == TM_IO_BINARY_PARTITIONED) ok = put_u32(out.f, partitions ? partitions[i] : 0u);
			ok = ok && put_filter(out.f, w) && put_u32(out.f, ke - kb);
			for (uint32_t x = kb; ok && x < ke; ++x) ok = put_u32(out.f, keys[x]);
		}
		if (!ok) return tmx::set_error(TM_ERR_INVALID, "%s: write failed at entry %llu", path, (unsigned long long)i);
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
