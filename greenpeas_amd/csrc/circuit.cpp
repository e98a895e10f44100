// circuit.cpp -- stim circuit text parser (see circuit.hpp).
#include "circuit.hpp"

#include <vector>

#include "workers.hpp"

#include <algorithm>
#include <cctype>
#include <charconv>
#include <cmath>
#include <cstdlib>
#include <cstring>

namespace gpa {
namespace {

struct Parser {
  std::string_view text;
  size_t pos = 0;
  uint32_t line = 0;

  [[noreturn]] void fail(const std::string &msg) const {
    throw InvalidArgument("line " + std::to_string(line) + ": " + msg);
  }

  // Next non-empty logical line with comments stripped; false at the end.
  bool nextLine(std::string_view &out) {
    while (pos < text.size()) {
      size_t end = text.find('\n', pos);
      if (end == std::string_view::npos) end = text.size();
      std::string_view ln = text.substr(pos, end - pos);
      pos = end + 1;
      line++;
      size_t hash = ln.find('#');
      if (hash != std::string_view::npos) ln = ln.substr(0, hash);
      size_t a = 0, b = ln.size();
      while (a < b && std::isspace(static_cast<unsigned char>(ln[a]))) a++;
      while (b > a && std::isspace(static_cast<unsigned char>(ln[b - 1]))) b--;
      if (a == b) continue;
      out = ln.substr(a, b - a);
      return true;
    }
    return false;
  }

  static bool iequals(std::string_view a, const char *b) {
    size_t n = std::strlen(b);
    if (a.size() != n) return false;
    for (size_t i = 0; i < n; ++i)
      if (std::toupper(static_cast<unsigned char>(a[i])) != b[i]) return false;
    return true;
  }

  void block(Circuit &c, int depth) {
    std::string_view ln;
    while (nextLine(ln)) {
      if (ln[0] == '}') {
        if (depth == 0) fail("unmatched '}'");
        return;
      }
      instruction(c, ln, depth);
    }
    if (depth) fail("unterminated REPEAT block");
  }

  static void appendBody(Circuit &dst, const Circuit &body, uint64_t reps) {
    dst.ops.reserve(dst.ops.size() + body.ops.size() * reps);
    dst.targets.reserve(dst.targets.size() + body.targets.size() * reps);
    for (uint64_t r = 0; r < reps; ++r) {
      const uint32_t shift = static_cast<uint32_t>(dst.targets.size());
      dst.targets.insert(dst.targets.end(), body.targets.begin(), body.targets.end());
      for (Op o : body.ops) {
        o.t0 += shift;
        dst.ops.push_back(o);
      }
    }
  }

  void instruction(Circuit &c, std::string_view ln, int depth) {
    size_t i = 0;
    while (i < ln.size() && (std::isalnum(static_cast<unsigned char>(ln[i])) || ln[i] == '_')) i++;
    if (i == 0) fail("cannot parse '" + std::string(ln) + "'");
    std::string_view name = ln.substr(0, i);
    std::string_view rest = ln.substr(i);
    double args[8];
    int nargs = 0;
    size_t k = 0;
    while (k < rest.size() && (rest[k] == ' ' || rest[k] == '\t')) k++;
    if (k < rest.size() && rest[k] == '(') {
      size_t close = rest.find(')', k);
      if (close == std::string_view::npos) fail("missing ')'");
      std::string inner(rest.substr(k + 1, close - k - 1));
      const char *p = inner.c_str();
      while (*p) {
        while (std::isspace(static_cast<unsigned char>(*p)) || *p == ',') p++;
        if (!*p) break;
        char *e;
        double v = std::strtod(p, &e);
        if (e == p) fail("bad argument '" + inner + "'");
        if (nargs == 8) fail("too many arguments");
        args[nargs++] = v;
        p = e;
      }
      rest = rest.substr(close + 1);
    } else {
      rest = rest.substr(k);
    }

    if (iequals(name, "REPEAT")) {
      size_t a = 0;
      while (a < rest.size() && std::isspace(static_cast<unsigned char>(rest[a]))) a++;
      uint64_t reps = 0;
      auto [ptr, ec] = std::from_chars(rest.data() + a, rest.data() + rest.size(), reps);
      if (ec != std::errc() || ptr == rest.data() + a) fail("bad REPEAT count");
      std::string_view tail = rest.substr(ptr - rest.data());
      size_t b = 0;
      while (b < tail.size() && std::isspace(static_cast<unsigned char>(tail[b]))) b++;
      if (b >= tail.size() || tail[b] != '{') fail("REPEAT needs '{'");
      Circuit body;
      block(body, depth + 1);
      appendBody(c, body, reps);
      return;
    }

    OpKind kinds[2];
    int nk = 1;
    if (iequals(name, "TICK")) kinds[0] = OpKind::Tick;
    else if (iequals(name, "R") || iequals(name, "RZ")) kinds[0] = OpKind::R;
    else if (iequals(name, "M") || iequals(name, "MZ")) kinds[0] = OpKind::M;
    else if (iequals(name, "MR") || iequals(name, "MRZ")) {
      kinds[0] = OpKind::M;
      kinds[1] = OpKind::R;
      nk = 2;
    } else if (iequals(name, "CX") || iequals(name, "CNOT") || iequals(name, "ZCX"))
      kinds[0] = OpKind::CX;
    else if (iequals(name, "H") || iequals(name, "H_XZ")) kinds[0] = OpKind::H;
    else if (iequals(name, "DEPOLARIZE1")) kinds[0] = OpKind::Dep1;
    else if (iequals(name, "DEPOLARIZE2")) kinds[0] = OpKind::Dep2;
    else if (iequals(name, "X_ERROR")) kinds[0] = OpKind::XErr;
    else if (iequals(name, "DETECTOR")) kinds[0] = OpKind::Detector;
    else if (iequals(name, "OBSERVABLE_INCLUDE")) kinds[0] = OpKind::Observable;
    else if (iequals(name, "QUBIT_COORDS") || iequals(name, "SHIFT_COORDS")) return;
    else throw InvalidArgument("Unsupported operation: " + std::string(name));

    const bool recTargets = kinds[0] == OpKind::Detector || kinds[0] == OpKind::Observable;
    const uint32_t t0 = static_cast<uint32_t>(c.targets.size());
    const char *p = rest.data(), *const pe = rest.data() + rest.size();
    auto space = [](char ch) { return ch == ' ' || ch == '\t' || ch == '\r' || ch == '\v' || ch == '\f'; };
    while (p < pe) {
      while (p < pe && space(*p)) p++;
      if (p >= pe) break;
      const char *e = p;
      while (e < pe && !space(*e)) e++;
      // decimal digits [d, de) of the token: all of it, or between "rec[-" and "]"
      const char *d = p, *de = e;
      if (recTargets) {
        if (e - p < 7 || !(p[0] == 'r' || p[0] == 'R') || std::string_view(p + 1, 4) != "ec[-" || e[-1] != ']')
          fail("bad record target '" + std::string(p, e) + "'");
        d = p + 5;
        de = e - 1;
      }
      uint64_t v = 0;
      const bool ok = de > d && de - d <= 10 && std::all_of(d, de, [](char ch) { return ch >= '0' && ch <= '9'; });
      if (ok)
        for (const char *q = d; q < de; ++q) v = v * 10 + static_cast<uint64_t>(*q - '0');
      if (!ok || v > 0x7FFFFFFFu || (recTargets && v == 0))
        fail(std::string(recTargets ? "bad record target '" : "bad target '") + std::string(p, e) + "'");
      c.targets.push_back(recTargets ? -static_cast<int32_t>(v) : static_cast<int32_t>(v));
      p = e;
    }
    const uint32_t nt = static_cast<uint32_t>(c.targets.size()) - t0;
    for (int j = 0; j < nk; ++j) {
      Op o{kinds[j], 0.0, t0, nt};
      switch (o.kind) {
      case OpKind::M:
        if (nargs > 1) fail("M takes at most one argument");
        o.arg = nargs ? args[0] : 0.0;
        break;
      case OpKind::Dep1:
      case OpKind::Dep2:
      case OpKind::XErr:
        if (nargs != 1) fail(std::string(name) + " needs one probability");
        o.arg = args[0];
        break;
      case OpKind::Observable:
        if (nargs != 1 || args[0] < 0 || args[0] != std::floor(args[0]) || args[0] > 1e9)
          fail("OBSERVABLE_INCLUDE needs a non-negative integer index");
        o.arg = args[0];
        break;
      case OpKind::Tick:
        if (nt) fail("TICK takes no targets");
        break;
      default:
        break;
      }
      if ((o.kind == OpKind::M || o.kind == OpKind::Dep1 || o.kind == OpKind::Dep2 ||
           o.kind == OpKind::XErr) &&
          !(o.arg >= 0.0 && o.arg <= 1.0))
        fail("probability out of range");
      if ((o.kind == OpKind::CX || o.kind == OpKind::Dep2) && (nt % 2))
        fail(std::string(name) + " needs target pairs");
      if (j == 1) {  // R half of MR: same targets
        o.t0 = static_cast<uint32_t>(c.targets.size());
        for (uint32_t q = 0; q < nt; ++q) c.targets.push_back(c.targets[t0 + q]);
      }
      c.ops.push_back(o);
    }
  }
};

}  // namespace

// Texts without REPEAT blocks above this size are parsed in line-aligned
// chunks on the worker pool and joined in order (same ops, targets and error
// messages as one pass).
constexpr size_t kParallelParseMinBytes = 65536;

Circuit Circuit::parse(std::string_view text) {
  Circuit c;
  const bool chunked = text.size() >= kParallelParseMinBytes && text.find('{') == std::string_view::npos &&
                       parallelWidth() > 1;
  if (!chunked) {
    Parser p{text};
    c.targets.reserve(text.size() / 3);
    p.block(c, 0);
  } else {
    const uint32_t parts = std::min<uint32_t>(4 * parallelWidth(), static_cast<uint32_t>(text.size() / 16384));
    std::vector<size_t> cut(parts + 1, text.size());
    cut[0] = 0;
    for (uint32_t k = 1; k < parts; ++k) {
      size_t at = std::max(cut[k - 1], text.size() * k / parts);
      const size_t nl = text.find('\n', at);
      cut[k] = nl == std::string_view::npos ? text.size() : nl + 1;
    }
    // each chunk's first line number
    std::vector<uint32_t> lines(parts + 1, 0);
    parallelFor(parts, [&](uint32_t k) {
      lines[k + 1] = static_cast<uint32_t>(std::count(text.begin() + cut[k], text.begin() + cut[k + 1], '\n'));
    });
    for (uint32_t k = 0; k < parts; ++k) lines[k + 1] += lines[k];
    std::vector<Circuit> pieces(parts);
    parallelFor(parts, [&](uint32_t k) {
      Parser p{text.substr(cut[k], cut[k + 1] - cut[k])};
      p.line = lines[k];
      pieces[k].targets.reserve((cut[k + 1] - cut[k]) / 3);
      p.block(pieces[k], 0);
    });
    size_t nOps = 0, nTargets = 0;
    for (const Circuit &pc : pieces) {
      nOps += pc.ops.size();
      nTargets += pc.targets.size();
    }
    c.ops.reserve(nOps);
    c.targets.resize(nTargets);
    std::vector<size_t> tOff(parts + 1, 0), oOff(parts + 1, 0);
    for (uint32_t k = 0; k < parts; ++k) {
      tOff[k + 1] = tOff[k] + pieces[k].targets.size();
      oOff[k + 1] = oOff[k] + pieces[k].ops.size();
    }
    if (nTargets >= (1ull << 32)) throw InvalidArgument("circuit too large: more than 2^32 targets");
    c.ops.resize(nOps);
    parallelFor(parts, [&](uint32_t k) {
      std::copy(pieces[k].targets.begin(), pieces[k].targets.end(), c.targets.begin() + tOff[k]);
      for (size_t i = 0; i < pieces[k].ops.size(); ++i) {
        Op o = pieces[k].ops[i];
        o.t0 += static_cast<uint32_t>(tOff[k]);
        c.ops[oOff[k] + i] = o;
      }
    });
  }
  // Stats, as stim's compute_stats feeds CircuitParameters
  // (Circuit.hpp:68-79).
  int64_t maxQubit = -1;
  uint32_t meas = 0;
  for (const Op &o : c.ops) {
    const int32_t *t = c.targets.data() + o.t0;
    switch (o.kind) {
    case OpKind::Tick:
      c.numTicks++;
      break;
    case OpKind::Detector:
    case OpKind::Observable:
      for (uint32_t j = 0; j < o.nt; ++j)
        if (static_cast<uint32_t>(-t[j]) > meas)
          throw InvalidArgument("record target rec[" + std::to_string(t[j]) +
                                "] refers before the first measurement");
      if (o.kind == OpKind::Detector) c.numDetectors++;
      else c.numObservables = std::max<uint32_t>(c.numObservables, static_cast<uint32_t>(o.arg) + 1);
      break;
    default:
      for (uint32_t j = 0; j < o.nt; ++j) maxQubit = std::max<int64_t>(maxQubit, t[j]);
      if (o.kind == OpKind::M) meas += o.nt;
      if (o.kind == OpKind::M || o.kind == OpKind::XErr || o.kind == OpKind::Dep1) c.numLocations += o.nt;
      if (o.kind == OpKind::Dep2) c.numLocations += o.nt / 2;
      break;
    }
  }
  c.numQubits = static_cast<uint32_t>(maxQubit + 1);
  c.numMeasurements = meas;
  return c;
}

namespace {

bool startsWith(std::string_view s, std::string_view p) { return s.substr(0, p.size()) == p; }

// Largest decimal number in a target token (qubit targets: "5", "!5",
// "X5*Y6" style products); record and sweep targets are skipped.
void maxQubitOf(std::string_view tok, uint64_t &maxq, bool &any) {
  if (startsWith(tok, "rec[") || startsWith(tok, "sweep[")) return;
  uint64_t v = 0;
  bool in = false;
  for (size_t i = 0; i <= tok.size(); ++i) {
    const char c = i < tok.size() ? tok[i] : ' ';
    if (c >= '0' && c <= '9') {
      v = v * 10 + static_cast<uint64_t>(c - '0');
      in = true;
    } else {
      if (in && (!any || v > maxq)) {
        maxq = v;
        any = true;
      }
      v = 0;
      in = false;
    }
  }
}

}  // namespace

CircuitSummary summarize(std::string_view text) {
  CircuitSummary r;
  std::vector<uint64_t> mult{1};
  uint64_t maxq = 0, ticks = 0;
  bool anyQ = false;
  size_t pos = 0;
  while (pos < text.size()) {
    size_t end = text.find('\n', pos);
    if (end == std::string_view::npos) end = text.size();
    std::string_view line = text.substr(pos, end - pos);
    pos = end + 1;
    const size_t hash = line.find('#');
    if (hash != std::string_view::npos) line = line.substr(0, hash);
    size_t b = 0;
    while (b < line.size() && (line[b] == ' ' || line[b] == '\t' || line[b] == '\r')) ++b;
    line = line.substr(b);
    if (line.empty()) continue;
    const uint64_t m = mult.back();
    if (line[0] == '}') {
      if (mult.size() > 1) mult.pop_back();
      continue;
    }
    size_t n = 0;
    while (n < line.size() && (std::isalnum(static_cast<unsigned char>(line[n])) || line[n] == '_')) ++n;
    const std::string_view name = line.substr(0, n);
    std::string_view rest = line.substr(n);
    if (name == "REPEAT") {
      uint64_t k = 0;
      for (char c : rest)
        if (c >= '0' && c <= '9') k = k * 10 + static_cast<uint64_t>(c - '0');
      mult.push_back(m * k);
      continue;
    }
    if (name == "TICK") {
      ticks += m;
      continue;
    }
    if (name == "DETECTOR") {
      r.detectors += m;
      continue;
    }
    if (name == "OBSERVABLE_INCLUDE") {
      uint64_t k = 0;
      const size_t open = rest.find('('), close = rest.find(')');
      if (open != std::string_view::npos && close != std::string_view::npos)
        for (size_t i = open + 1; i < close; ++i)
          if (rest[i] >= '0' && rest[i] <= '9') k = k * 10 + static_cast<uint64_t>(rest[i] - '0');
      r.observables = std::max(r.observables, k + 1);
      continue;
    }
    if (name == "QUBIT_COORDS" || name == "SHIFT_COORDS") continue;
    if (!rest.empty() && rest[0] == '(') {
      const size_t close = rest.find(')');
      rest = close == std::string_view::npos ? std::string_view() : rest.substr(close + 1);
    }
    size_t i = 0;
    while (i < rest.size()) {
      while (i < rest.size() && (rest[i] == ' ' || rest[i] == '\t' || rest[i] == '\r')) ++i;
      size_t j = i;
      while (j < rest.size() && rest[j] != ' ' && rest[j] != '\t' && rest[j] != '\r') ++j;
      if (j > i) maxQubitOf(rest.substr(i, j - i), maxq, anyQ);
      i = j;
    }
  }
  r.layers = ticks + 1;
  r.qubits = anyQ ? maxq + 1 : 0;
  return r;
}

}  // namespace gpa
