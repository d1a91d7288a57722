// fastx_input.hpp — the input stage of dnaln_map_file: FASTA/FASTQ (plain or gzip) cut into
// batches and parsed by several threads when the file is plain 4-line FASTQ, by the reference's
// sequential rules (src/seqio.cpp:59-87, as the record machine of fastx_record.hpp) otherwise.
//
// A 4-line FASTQ file can be cut at any record start, and record starts can be found from any
// point: a line that begins with '@' and whose line two further on begins with '+' is a header
// (a quality line that begins with '@' is followed by a header and then a sequence line, which
// in this form never begins with '+'). Each part must then parse exactly up to the next part's
// start; if any record falls outside the form whose result the sequential reader is known to
// give, the batch is parsed again by the sequential reader.
#pragma once

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>
#include <zlib.h>

#include <algorithm>
#include <cstring>
#include <exception>
#include <memory>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "dnaln_hip.h"
#include "host/fastx_record.hpp"

namespace dnaln_amd {
namespace fastx {

struct Failure : std::runtime_error {
  int code;
  Failure(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// nt4 of src/common.hpp:23-31 as a table.
struct Nt4 {
  uint8_t t[256];
  Nt4() {
    std::memset(t, 4, sizeof t);
    t[(unsigned char)'A'] = t[(unsigned char)'a'] = 0;
    t[(unsigned char)'C'] = t[(unsigned char)'c'] = 1;
    t[(unsigned char)'G'] = t[(unsigned char)'g'] = 2;
    t[(unsigned char)'T'] = t[(unsigned char)'t'] = t[(unsigned char)'U'] = t[(unsigned char)'u'] = 3;
  }
};
const Nt4 kNt4;

// Grow-only byte buffer without zero-filling.
struct Bytes {
  std::unique_ptr<char[]> p;
  size_t n = 0, cap = 0;
  void reserve(size_t want) {
    if (want <= cap) return;
    const size_t c = std::max(want, cap + cap / 2);
    std::unique_ptr<char[]> q(new char[c]);
    if (n) std::memcpy(q.get(), p.get(), n);
    p.swap(q);
    cap = c;
  }
  void append(const char* s, size_t k) {
    reserve(n + k);
    std::memcpy(p.get() + n, s, k);
    n += k;
  }
  char* data() { return p.get(); }
  const char* data() const { return p.get(); }
};

// Reads of one parser thread: names, sequences and qualities back to back.
struct Part {
  Bytes names, seq, qual;
  std::vector<uint64_t> name_off{0}, seq_off{0};  // n + 1 entries
  std::vector<uint8_t> fq;                         // 1: FASTQ record (has a quality string)
  size_t n() const { return fq.size(); }
  void clear() {
    names.n = seq.n = qual.n = 0;
    name_off.assign(1, 0);
    seq_off.assign(1, 0);
    fq.clear();
  }
  void add(const char* name, size_t nl, const char* s, size_t sl, const char* q, bool fastq) {
    names.append(name, nl);
    name_off.push_back(names.n);
    seq.append(s, sl);
    seq_off.push_back(seq.n);
    if (fastq) qual.append(q, sl);
    else qual.reserve(qual.n + sl), qual.n += sl;  // FASTA: no quality (bytes unused)
    fq.push_back(fastq ? 1 : 0);
  }
};

// The input stream: bytes [pos, buf.n) are unparsed.
struct Input {
  gzFile f = nullptr;            // gzip input (or standard input), read into buf
  const char* map = nullptr;     // a plain file, mapped whole: parsed in place
  size_t map_len = 0;
  Bytes buf;
  size_t pos = 0;                // bytes [pos, size()) are unparsed
  bool eof = false;              // everything up to size() is all there is
  ~Input() {
    if (f) gzclose(f);
    if (map) munmap(const_cast<char*>(map), map_len);
  }
  const char* data() const { return map ? map : buf.data(); }
  size_t size() const { return map ? map_len : buf.n; }
  // Opens `path` ("-": standard input): a plain file is mapped, anything else read through zlib.
  void open(const char* path) {
    const bool in_std = std::strcmp(path, "-") == 0;
    const int fd = in_std ? dup(0) : ::open(path, O_RDONLY);
    if (fd < 0) throw Failure(DNALN_E_IO, std::string("cannot open '") + path + "'");
    struct stat sb;
    unsigned char magic[2] = {0, 0};
    if (!in_std && fstat(fd, &sb) == 0 && S_ISREG(sb.st_mode) &&
        (sb.st_size < 2 || (pread(fd, magic, 2, 0) == 2 && !(magic[0] == 0x1f && magic[1] == 0x8b)))) {
      map_len = (size_t)sb.st_size;
      if (map_len) {
        void* m = mmap(nullptr, map_len, PROT_READ, MAP_PRIVATE, fd, 0);
        if (m == MAP_FAILED) {
          ::close(fd);
          throw Failure(DNALN_E_IO, std::string("cannot map '") + path + "'");
        }
        madvise(m, map_len, MADV_SEQUENTIAL);
        map = static_cast<const char*>(m);
      }
      ::close(fd);
      eof = true;
      return;
    }
    f = gzdopen(fd, "rb");
    if (!f) {
      ::close(fd);
      throw Failure(DNALN_E_IO, std::string("cannot open '") + path + "'");
    }
    gzbuffer(f, 1 << 20);
  }
  // Drops the parsed prefix and reads until `want` unparsed bytes are held or the input ends (a
  // mapped file holds everything already).
  void top_up(size_t want) {
    if (map || !f) return;
    if (pos) {
      std::memmove(buf.data(), buf.data() + pos, buf.n - pos);
      buf.n -= pos;
      pos = 0;
    }
    while (!eof && buf.n < want) {
      const size_t chunk = std::min<size_t>(std::max<size_t>(want - buf.n, 1 << 20), 1u << 30);
      buf.reserve(buf.n + chunk);
      const int got = gzread(f, buf.data() + buf.n, (unsigned)chunk);
      if (got < 0) throw Failure(DNALN_E_IO, "read error");
      if (got == 0) eof = true;
      buf.n += (size_t)got;
    }
  }
  // The line source of the record machine (host/fastx_record.hpp): the next whole line of the
  // unparsed bytes, reading more as needed; unread() puts the last line back.
  bool next(std::string_view& ln) {
    for (;;) {
      const char* p = data() + pos;
      if (cut_line(p, data() + size(), eof, ln)) {
        last = pos;
        pos = (size_t)(p - data());
        return true;
      }
      if (eof) return false;
      top_up(size() - pos + (1 << 20));
    }
  }
  void unread() { pos = last; }
  size_t last = 0;  // start of the line next() gave last
};

// One record by the sequential rules into `part`; false at the end of the input.
inline bool next_general(Input& in, Part& part, RecordMachine& m) {
  try {
    if (!read_record(in, m)) return false;
  } catch (const RecordError& e) {
    throw Failure(DNALN_E_IO, e.what());
  }
  part.add(m.name.data(), m.name.size(), m.seq.data(), m.seq.size(), m.qual.data(), m.fastq);
  return true;
}

// A line [b, e) of the text (without '\n' and a trailing '\r'); `next` is the start of the line
// after it. False when no '\n' ends it before `end` (unless `final`, where `end` ends it).
struct Line {
  const char *b, *e, *next;
};
inline bool get_line(const char* p, const char* end, bool final, Line& l) {
  const char* nl = static_cast<const char*>(std::memchr(p, '\n', (size_t)(end - p)));
  if (!nl) {
    if (!final || p >= end) return false;
    nl = end;
  }
  l.b = p;
  l.e = nl > p && nl[-1] == '\r' ? nl - 1 : nl;
  l.next = nl < end ? nl + 1 : end;
  return true;
}

// Four-line FASTQ records from p (a header line) up to `stop`. Records must tile [p, stop)
// exactly unless `last`, where parsing ends at the first record not complete before `end`
// (`final`: the input ends at `end`). Returns where parsing ended, or nullptr when a record
// is outside the form whose result the reference's reader is known to match: header '@';
// sequence line not starting with '@', '+' or '>'; a '+' line; a quality line as long as the
// sequence (empty for an empty sequence); and the next record starting right after.
inline const char* parse_fastq4(const char* p, const char* stop, const char* end, bool last, bool final, Part& part) {
  while (p < stop) {
    Line h, s, plus, q;
    if (*p != '@') return nullptr;
    if (!get_line(p, end, final, h) || !get_line(h.next, end, final, s) || !get_line(s.next, end, final, plus) ||
        !get_line(plus.next, end, final, q)) {
      if (last) return p;  // incomplete: the caller keeps it for the next batch
      return nullptr;
    }
    if (s.e > s.b && (*s.b == '@' || *s.b == '+' || *s.b == '>')) return nullptr;
    if (plus.e == plus.b || *plus.b != '+') return nullptr;
    if (q.e - q.b != s.e - s.b) return nullptr;
    if (!last && q.next > stop) return nullptr;
    if (q.next < end && *q.next != '@') return nullptr;
    const char* ne = h.b + 1;
    while (ne < h.e && *ne != ' ' && *ne != '\t') ++ne;
    part.add(h.b + 1, (size_t)(ne - h.b - 1), s.b, (size_t)(s.e - s.b), q.b, true);
    p = q.next;
    if (!last && p == stop) return p;
  }
  return p;
}

// True when the line at p starts a record of the four-line form: '@' here and '+' two lines on.
inline bool fastq_header_at(const char* p, const char* end) {
  if (p >= end || *p != '@') return false;
  Line a, b;
  if (!get_line(p, end, false, a) || !get_line(a.next, end, false, b)) return false;
  return b.next < end && *b.next == '+';
}

template <class F>
inline void parallel(int threads, int n, F&& f) {
  if (n <= 1 || threads <= 1) {
    for (int i = 0; i < n; ++i) f(i);
    return;
  }
  std::vector<std::thread> th;
  std::vector<std::exception_ptr> err(n);
  for (int i = 0; i < n; ++i)
    th.emplace_back([&, i] {
      try {
        f(i);
      } catch (...) {
        err[i] = std::current_exception();
      }
    });
  for (auto& t : th) t.join();
  for (auto& e : err)
    if (e) std::rethrow_exception(e);
}

// Cuts the input into batches of about batch_bytes: plain 4-line FASTQ into `threads` parts
// parsed at once, anything else through the sequential reader (from the first batch that needs
// it to the end).
class Batcher {
 public:
  Batcher(const char* path, uint64_t batch_bytes, int threads) : batch_bytes(batch_bytes), threads(threads) {
    in.open(path);
  }
  // True while every batch so far took the parallel parser.
  bool parallel_fastq() const { return !general; }
  // Fills parts (resized to `threads`) with the next batch; false at the end of the input.
  bool next(std::vector<Part>& parts) {
    for (Part& p : parts) p.clear();
    if (parts.size() != (size_t)threads) parts.resize(threads);
    if (!general) {
      in.top_up(batch_bytes);
      const char* b = in.data() + in.pos;
      const char* e = in.data() + std::min<size_t>(in.size(), in.pos + batch_bytes);
      const bool final = in.eof && e == in.data() + in.size();
      if (b == e && final) return false;
      if (*b != '@') {
        general = true;
      } else {
        // starts of the pieces: at record starts found from even cuts
        std::vector<const char*> cut{b};
        for (int t = 1; t < threads; ++t) {
          const char* q = b + (size_t)(e - b) * t / threads;
          const char* nl = static_cast<const char*>(std::memchr(q, '\n', (size_t)(e - q)));
          if (!nl) break;
          const char* l = nl + 1;
          for (int tries = 0; tries < 8 && l < e && !fastq_header_at(l, e); ++tries) {
            const char* n2 = static_cast<const char*>(std::memchr(l, '\n', (size_t)(e - l)));
            l = n2 ? n2 + 1 : e;
          }
          if (l < e && fastq_header_at(l, e) && l > cut.back()) cut.push_back(l);
        }
        const int np = (int)cut.size();
        cut.push_back(e);
        std::vector<const char*> ends(np, nullptr);
        parallel(threads, np, [&](int t) {
          ends[t] = parse_fastq4(cut[t], cut[t + 1], e, t == np - 1, final, parts[t]);
        });
        bool ok = true;
        for (int t = 0; t < np; ++t) ok = ok && ends[t] != nullptr;
        if (ok && ends[np - 1] == b) ok = false;  // not one complete record in the batch
        if (ok) {
          in.pos += (size_t)(ends[np - 1] - b);
        } else {
          for (Part& p : parts) p.clear();
          general = true;
        }
      }
    }
    if (general) {
      Part& p = parts[0];
      while (p.seq.n + p.names.n < batch_bytes / 2 && next_general(in, p, machine)) {
      }
      if (p.n() == 0) return false;
    }
    return true;
  }


 private:
  uint64_t batch_bytes;
  int threads;
  Input in;
  RecordMachine machine;  // the sequential rules, once a batch needs them
  bool general = false;
};

}  // namespace fastx
}  // namespace dnaln_amd
