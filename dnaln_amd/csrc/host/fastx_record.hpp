// fastx_record.hpp — the FASTA/FASTQ record rules as a line-fed state machine, shared by the
// sequential readers (SeqReader, and the batcher's fallback in fastx_input.hpp).
//
// The rules are those of the reference reader (src/seqio.cpp:59-87):
//   1. Input is cut into lines at '\n'; one trailing '\r' is dropped from each line. Bytes after
//      the last '\n' form a final line when there are any.
//   2. While no record is open, lines that do not begin with '>' or '@' are ignored.
//   3. A line beginning with '>' (FASTA) or '@' (FASTQ) opens a record. Its name is the text after
//      the marker up to the first space or tab.
//   4. Every following line that does not begin with '>', '@' or '+' is appended to the sequence
//      as it is (empty lines add nothing).
//   5. FASTA: a line beginning with '>' or '@' closes the record and opens the next one; a line
//      beginning with '+' closes it and is then ignored (rule 2); the end of input closes it.
//   6. FASTQ: the sequence must end at a line beginning with '+' (a header line or the end of input
//      there is the error "malformed FASTQ record '<name>'"). That line is skipped, then whole
//      lines are appended to the quality while it is shorter than the sequence; it must come out
//      exactly as long ("FASTQ record '<name>': quality length mismatch"). An empty sequence takes
//      no quality line.
// The machine only ever looks at the line it is given, so any line source can drive it.
#pragma once

#include <cstring>
#include <stdexcept>
#include <string>
#include <string_view>

namespace dnaln_amd {
namespace fastx {

// Error of a malformed record; the message is the reference's (without its "[dnaln] error: ").
struct RecordError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

class RecordMachine {
 public:
  // What a fed line did: nothing to hand out yet, a record is complete (the line is used up), or a
  // record is complete and the same line must be fed again (it opens the next record).
  enum Step { kMore, kRecord, kRecordRefeed };

  std::string name, seq, qual;  // the record being built (complete after kRecord / kRecordRefeed)
  bool fastq = false;

  Step feed(std::string_view ln) {
    const char lead = ln.empty() ? '\0' : ln.front();
    const bool header = lead == '>' || lead == '@';
    switch (state_) {
      case kIdle:
        if (header) open(ln);
        return kMore;
      case kSequence:
        if (header) {
          if (fastq) throw RecordError("malformed FASTQ record '" + name + "'");
          state_ = kIdle;
          return kRecordRefeed;
        }
        if (lead != '+') {
          seq.append(ln);
          return kMore;
        }
        if (!fastq || seq.empty()) {  // FASTA ends here; an empty FASTQ sequence has no quality
          state_ = kIdle;
          return kRecord;
        }
        state_ = kQuality;
        return kMore;
      case kQuality:
        qual.append(ln);
        if (qual.size() < seq.size()) return kMore;
        state_ = kIdle;
        if (qual.size() != seq.size()) throw RecordError("FASTQ record '" + name + "': quality length mismatch");
        return kRecord;
    }
    return kMore;
  }

  // End of input: true when it completes an open record.
  bool finish() {
    const State s = state_;
    state_ = kIdle;
    if (s == kIdle) return false;
    if (s == kQuality) throw RecordError("FASTQ record '" + name + "': quality length mismatch");
    if (fastq) throw RecordError("malformed FASTQ record '" + name + "'");
    return true;
  }

 private:
  enum State { kIdle, kSequence, kQuality };
  State state_ = kIdle;

  void open(std::string_view ln) {
    fastq = ln.front() == '@';
    size_t e = 1;
    while (e < ln.size() && ln[e] != ' ' && ln[e] != '\t') ++e;
    name.assign(ln.data() + 1, e - 1);
    seq.clear();
    qual.clear();
    state_ = kSequence;
  }
};

// Cuts a byte range into lines by rule 1: [p, end) holds whole lines when `final`; otherwise a
// line needs its '\n' before `end`. On success `ln` is the line (no '\n', no trailing '\r') and
// `p` moves past it.
inline bool cut_line(const char*& p, const char* end, bool final, std::string_view& ln) {
  if (p >= end) return false;
  const char* nl = static_cast<const char*>(std::memchr(p, '\n', (size_t)(end - p)));
  const char* stop = nl ? nl : end;
  if (!nl && !final) return false;
  size_t len = (size_t)(stop - p);
  if (len && p[len - 1] == '\r') --len;
  ln = std::string_view(p, len);
  p = nl ? nl + 1 : end;
  return true;
}

// Runs the machine over a line source until one record is complete; false at the end of the
// input. The source gives lines with next(std::string_view&) and takes the last one back with
// unread(), so a line that closes one record and opens the next is read twice.
template <class Source>
bool read_record(Source& src, RecordMachine& m) {
  std::string_view ln;
  while (src.next(ln)) {
    const RecordMachine::Step s = m.feed(ln);
    if (s == RecordMachine::kMore) continue;
    if (s == RecordMachine::kRecordRefeed) src.unread();
    return true;
  }
  return m.finish();
}

}  // namespace fastx
}  // namespace dnaln_amd
