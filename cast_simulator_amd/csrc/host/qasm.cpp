// OpenQASM 2.0 ingest: lexer, recursive-descent parser with exact
// rational-multiple-of-pi angles, serialisation and conversion to gates.
#include "qasm.hpp"

#include <charconv>
#include <cmath>
#include <cstdlib>
#include <limits>
#include <memory>
#include <numeric>

namespace cast {

// ── Fraction ───────────────────────────────────────────────────────────────

Fraction Fraction::make(int64_t n, int64_t d) {
  if (d == 0)
    throw Error("division by zero in angle expression");
  if (n == 0)
    return Fraction{0, 1};
  if (d < 0) {
    n = -n;
    d = -d;
  }
  const int64_t g = std::gcd(n < 0 ? -n : n, d);
  n /= g;
  d /= g;
  if (n > std::numeric_limits<int32_t>::max() || n < std::numeric_limits<int32_t>::min() ||
      d > std::numeric_limits<int32_t>::max())
    throw Error("angle fraction does not fit in 32 bits");
  return Fraction{int32_t(n), int32_t(d)};
}

namespace {

Fraction f_add(Fraction a, Fraction b) {
  return Fraction::make(int64_t(a.num) * b.den + int64_t(b.num) * a.den, int64_t(a.den) * b.den);
}
Fraction f_sub(Fraction a, Fraction b) {
  return Fraction::make(int64_t(a.num) * b.den - int64_t(b.num) * a.den, int64_t(a.den) * b.den);
}
Fraction f_mul(Fraction a, Fraction b) {
  return Fraction::make(int64_t(a.num) * b.num, int64_t(a.den) * b.den);
}
Fraction f_div(Fraction a, Fraction b) {
  if (b.num == 0)
    throw Error("division by zero in angle expression");
  return Fraction::make(int64_t(a.num) * b.den, int64_t(a.den) * b.num);
}

std::string shortest(double v) {
  // Shortest round-trip digits, positional (never exponent) notation.
  char buf[400];
  const auto r = std::to_chars(buf, buf + sizeof(buf), v, std::chars_format::fixed);
  return std::string(buf, r.ptr);
}

// ── Lexer ──────────────────────────────────────────────────────────────────

enum class Tok {
  Ident, Int, Float, Pi, LParen, RParen, LBracket, RBracket, Comma, Semi,
  Plus, Minus, Star, Slash, Caret, Equal, Other, End
};

struct Token {
  Tok kind = Tok::End;
  std::string text; // identifier or the unsupported character
  int64_t i = 0;
  double f = 0.0;
};

std::string describe(const Token &t) {
  switch (t.kind) {
  case Tok::Ident: return "identifier `" + t.text + "`";
  case Tok::Int: return "integer `" + std::to_string(t.i) + "`";
  case Tok::Float: return "float `" + shortest(t.f) + "`";
  case Tok::Pi: return "`pi`";
  case Tok::LParen: return "`(`";
  case Tok::RParen: return "`)`";
  case Tok::LBracket: return "`[`";
  case Tok::RBracket: return "`]`";
  case Tok::Comma: return "`,`";
  case Tok::Semi: return "`;`";
  case Tok::Plus: return "`+`";
  case Tok::Minus: return "`-`";
  case Tok::Star: return "`*`";
  case Tok::Slash: return "`/`";
  case Tok::Caret: return "`^`";
  case Tok::Equal: return "`=`";
  case Tok::Other: return "unsupported character `" + t.text + "`";
  case Tok::End: return "<end-of-input>";
  }
  return "?";
}

bool ident_start(char c) { return std::isalpha(static_cast<unsigned char>(c)) || c == '_'; }
bool ident_char(char c) { return std::isalnum(static_cast<unsigned char>(c)) || c == '_'; }
bool digit(char c) { return c >= '0' && c <= '9'; }

std::vector<Token> lex(const std::string &s) {
  std::vector<Token> out;
  size_t i = 0;
  const size_t n = s.size();
  auto single = [&](Tok k) {
    Token t;
    t.kind = k;
    out.push_back(t);
    ++i;
  };
  while (i < n) {
    const char c = s[i];
    if (c == ' ' || c == '\t' || c == '\n' || c == '\r') {
      ++i;
    } else if (c == '/' && i + 1 < n && s[i + 1] == '/') {
      while (i < n && s[i] != '\n')
        ++i;
    } else if (digit(c) || (c == '.' && i + 1 < n && digit(s[i + 1]))) {
      // digits, at most one '.', then an optional exponent e[+-]digits
      size_t j = i;
      bool is_float = false;
      while (j < n && (digit(s[j]) || (s[j] == '.' && !is_float))) {
        is_float = is_float || s[j] == '.';
        ++j;
      }
      if (j < n && (s[j] == 'e' || s[j] == 'E')) {
        size_t k = j + 1;
        if (k < n && (s[k] == '+' || s[k] == '-'))
          ++k;
        if (k >= n || !digit(s[k]))
          throw Error("invalid scientific literal `" + s.substr(i, j - i) + "`");
        while (k < n && digit(s[k]))
          ++k;
        j = k;
        is_float = true;
      }
      const std::string lit = s.substr(i, j - i);
      Token t;
      if (is_float) {
        t.kind = Tok::Float;
        char *end = nullptr;
        t.f = std::strtod(lit.c_str(), &end);
      } else {
        t.kind = Tok::Int;
        const auto r = std::from_chars(lit.data(), lit.data() + lit.size(), t.i);
        if (r.ec != std::errc())
          throw Error("invalid integer literal `" + lit + "`");
      }
      out.push_back(t);
      i = j;
    } else if (ident_start(c)) {
      size_t j = i;
      while (j < n && ident_char(s[j]))
        ++j;
      Token t;
      t.text = s.substr(i, j - i);
      t.kind = t.text == "pi" ? Tok::Pi : Tok::Ident;
      out.push_back(t);
      i = j;
    } else {
      switch (c) {
      case '(': single(Tok::LParen); break;
      case ')': single(Tok::RParen); break;
      case '[': single(Tok::LBracket); break;
      case ']': single(Tok::RBracket); break;
      case ',': single(Tok::Comma); break;
      case ';': single(Tok::Semi); break;
      case '+': single(Tok::Plus); break;
      case '-': single(Tok::Minus); break;
      case '*': single(Tok::Star); break;
      case '/': single(Tok::Slash); break;
      case '^': single(Tok::Caret); break;
      case '=': single(Tok::Equal); break;
      default: {
        Token t;
        t.kind = Tok::Other;
        t.text = std::string(1, c);
        out.push_back(t);
        ++i;
      }
      }
    }
  }
  out.push_back(Token{});
  return out;
}

// ── Angle expressions ──────────────────────────────────────────────────────
//
// Every node carries its float value and a symbolic state: Exact (a rational,
// pi_power 0, or a rational times pi, pi_power 1), Inexact (float literals,
// powers, pi * pi, mixed sums) or Error.  As in the reference's symbolic pass,
// operands are looked at left to right and an Inexact left operand settles
// the node as Inexact before the right one is looked at, so an error (a zero
// divisor, an integer beyond 32 bits) counts only where it is reached; an
// Error reaching the root fails the parse.

enum class Sym { Exact, Inexact, Error };

struct Value {
  double f = 0.0;
  Sym s = Sym::Inexact;
  int pi_power = 0;
  Fraction r;
  std::string err;
};

Value v_error(double f, std::string msg) {
  Value v;
  v.f = f;
  v.s = Sym::Error;
  v.err = std::move(msg);
  return v;
}

Value v_int(int64_t x) {
  Value v;
  v.f = double(x);
  if (x < std::numeric_limits<int32_t>::min() || x > std::numeric_limits<int32_t>::max())
    return v_error(v.f, "integer `" + std::to_string(x) + "` is too large");
  v.s = Sym::Exact;
  v.r = Fraction{int32_t(x), 1};
  return v;
}

Value binary(char op, const Value &a, const Value &b) {
  Value v;
  switch (op) {
  case '+': v.f = a.f + b.f; break;
  case '-': v.f = a.f - b.f; break;
  case '*': v.f = a.f * b.f; break;
  case '/': v.f = a.f / b.f; break;
  case '^': v.f = std::pow(a.f, b.f); break;
  }
  if (a.s != Sym::Exact)
    return a.s == Sym::Error ? v_error(v.f, a.err) : v;
  if (b.s != Sym::Exact)
    return b.s == Sym::Error ? v_error(v.f, b.err) : v;
  if (op == '^')
    return v;
  try {
    if (op == '+' || op == '-') {
      if (a.pi_power != b.pi_power)
        return v;
      v.pi_power = a.pi_power;
      v.r = op == '+' ? f_add(a.r, b.r) : f_sub(a.r, b.r);
    } else if (op == '*') {
      if (a.pi_power + b.pi_power > 1)
        return v;
      v.pi_power = a.pi_power + b.pi_power;
      v.r = f_mul(a.r, b.r);
    } else { // '/'
      if (b.pi_power > a.pi_power)
        return v; // rational / (rational * pi) is not a multiple of pi
      if (b.r.num == 0)
        return v_error(v.f, "division by zero in angle expression");
      v.pi_power = a.pi_power - b.pi_power;
      v.r = f_div(a.r, b.r);
    }
  } catch (const Error &e) {
    return v_error(v.f, e.what());
  }
  v.s = Sym::Exact;
  return v;
}

class Parser {
public:
  explicit Parser(std::vector<Token> toks) : t_(std::move(toks)) {}

  QasmCircuit circuit() {
    QasmCircuit c;
    while (peek().kind != Tok::End) {
      if (peek().kind == Tok::Semi) {
        next();
        continue;
      }
      if (declaration())
        continue;
      c.gates.push_back(gate());
    }
    return c;
  }

private:
  const Token &peek() const { return t_[std::min(p_, t_.size() - 1)]; }
  Token next() {
    Token t = peek();
    if (p_ + 1 < t_.size())
      ++p_;
    return t;
  }
  void expect(Tok k, const char *what) {
    if (peek().kind != k)
      throw Error(std::string("expected ") + what + ", found " + describe(peek()));
    next();
  }
  std::string ident() {
    if (peek().kind != Tok::Ident)
      throw Error("expected identifier, found " + describe(peek()));
    return next().text;
  }
  static bool iequals(const std::string &a, const char *b) {
    size_t i = 0;
    for (; i < a.size() && b[i]; ++i)
      if (std::tolower(static_cast<unsigned char>(a[i])) != b[i])
        return false;
    return i == a.size() && !b[i];
  }
  void skip_statement(const char *what) {
    while (peek().kind != Tok::Semi) {
      if (peek().kind == Tok::End)
        throw Error(std::string("expected `;` to terminate ") + what + ", found end-of-input");
      next();
    }
    next();
  }

  bool declaration() {
    if (peek().kind != Tok::Ident)
      return false;
    const std::string &kw = peek().text;
    if (iequals(kw, "openqasm")) {
      next();
      skip_statement("OPENQASM header");
      return true;
    }
    if (iequals(kw, "include")) {
      next();
      skip_statement("include statement");
      return true;
    }
    if (iequals(kw, "qreg")) {
      next();
      if (has_qreg_)
        throw Error("multiple qreg declarations are not supported in this parser");
      std::string name = ident();
      expect(Tok::LBracket, "`[`");
      if (peek().kind != Tok::Int)
        throw Error("expected integer qreg size, found " + describe(peek()));
      const uint32_t size = index(next().i);
      expect(Tok::RBracket, "`]`");
      expect(Tok::Semi, "`;`");
      has_qreg_ = true;
      qreg_name_ = std::move(name);
      qreg_size_ = size;
      return true;
    }
    return false;
  }

  static uint32_t index(int64_t v) {
    if (v < 0 || v > int64_t(std::numeric_limits<uint32_t>::max()))
      throw Error("qubit index must be a non-negative 32-bit integer, found `" +
                  std::to_string(v) + "`");
    return uint32_t(v);
  }
  uint32_t bounded(uint32_t q) const {
    if (has_qreg_ && q >= qreg_size_)
      throw Error("qubit index `" + std::to_string(q) + "` is out of bounds for qreg size `" +
                  std::to_string(qreg_size_) + "`");
    return q;
  }

  uint32_t qubit() {
    if (peek().kind == Tok::Int)
      return bounded(index(next().i));
    if (peek().kind != Tok::Ident)
      throw Error("expected qubit operand, found " + describe(peek()));
    const std::string name = next().text;
    if (peek().kind == Tok::LBracket) {
      next();
      if (peek().kind != Tok::Int)
        throw Error("expected integer qubit index, found " + describe(peek()));
      const uint32_t q = index(next().i);
      expect(Tok::RBracket, "`]`");
      if (has_qreg_ && name != qreg_name_)
        throw Error("unknown qreg `" + name + "`; expected `" + qreg_name_ + "`");
      return bounded(q);
    }
    size_t d = name.size();
    while (d > 0 && digit(name[d - 1]))
      --d;
    if (d == name.size())
      throw Error("expected qubit like `q[0]` or `q0`, found identifier `" + name + "`");
    uint64_t v = 0;
    const auto r = std::from_chars(name.data() + d, name.data() + name.size(), v);
    if (r.ec != std::errc() || v > std::numeric_limits<uint32_t>::max())
      throw Error("invalid trailing index in identifier `" + name + "`");
    return bounded(uint32_t(v));
  }

  // expr := term (('+'|'-') term)* ; term := unary (('*'|'/') unary)* ;
  // unary := '-' unary | power ; power := primary ('^' unary)?
  Value expr() {
    Value v = term();
    while (peek().kind == Tok::Plus || peek().kind == Tok::Minus) {
      const char op = next().kind == Tok::Plus ? '+' : '-';
      v = binary(op, v, term());
    }
    return v;
  }
  Value term() {
    Value v = unary();
    while (peek().kind == Tok::Star || peek().kind == Tok::Slash) {
      const char op = next().kind == Tok::Star ? '*' : '/';
      v = binary(op, v, unary());
    }
    return v;
  }
  Value unary() {
    if (peek().kind == Tok::Minus) {
      next();
      Value v = unary();
      v.f = -v.f;
      if (v.s == Sym::Exact)
        v.r = Fraction::make(-int64_t(v.r.num), v.r.den);
      return v;
    }
    Value base = primary();
    if (peek().kind == Tok::Caret) {
      next();
      return binary('^', base, unary());
    }
    return base;
  }
  Value primary() {
    const Token t = peek();
    switch (t.kind) {
    case Tok::Int:
      next();
      return v_int(t.i);
    case Tok::Float: {
      next();
      Value v;
      v.f = t.f;
      return v;
    }
    case Tok::Pi: {
      next();
      Value v;
      v.f = M_PI;
      v.s = Sym::Exact;
      v.pi_power = 1;
      v.r = Fraction{1, 1};
      return v;
    }
    case Tok::LParen: {
      next();
      Value v = expr();
      expect(Tok::RParen, "`)`");
      return v;
    }
    default:
      throw Error("expected expression term, found " + describe(t));
    }
  }
  static QasmAngle angle_of(const Value &v) {
    if (v.s == Sym::Error)
      throw Error(v.err);
    QasmAngle a;
    if (v.s == Sym::Exact && v.pi_power == 1) {
      a.pi_multiple = true;
      a.frac = v.r;
    } else {
      a.number = v.f;
    }
    return a;
  }
  QasmAngle paren_angle() {
    expect(Tok::LParen, "`(`");
    const Value v = expr();
    expect(Tok::RParen, "`)`");
    return angle_of(v);
  }

  QasmGate gate() {
    const std::string name = ident();
    QasmGate g;
    auto qubits = [&](int k) {
      for (int i = 0; i < k; ++i) {
        if (i)
          expect(Tok::Comma, "`,`");
        g.qubits.push_back(qubit());
      }
    };
    static const struct {
      const char *name;
      QasmOp op;
      int params, qubits;
    } table[] = {{"x", QasmOp::X, 0, 1},    {"y", QasmOp::Y, 0, 1},      {"z", QasmOp::Z, 0, 1},
                 {"h", QasmOp::H, 0, 1},    {"s", QasmOp::S, 0, 1},      {"t", QasmOp::T, 0, 1},
                 {"rx", QasmOp::RX, 1, 1},  {"ry", QasmOp::RY, 1, 1},    {"rz", QasmOp::RZ, 1, 1},
                 {"u3", QasmOp::U3, 3, 1},  {"cx", QasmOp::CX, 0, 2},    {"cz", QasmOp::CZ, 0, 2},
                 {"cp", QasmOp::CP, 1, 2},  {"swap", QasmOp::SWAP, 0, 2}, {"ccx", QasmOp::CCX, 0, 3}};
    for (const auto &e : table) {
      if (name != e.name)
        continue;
      g.op = e.op;
      if (e.params == 1) {
        g.params.push_back(paren_angle());
      } else if (e.params == 3) {
        expect(Tok::LParen, "`(`");
        for (int i = 0; i < 3; ++i) {
          if (i)
            expect(Tok::Comma, "`,`");
          g.params.push_back(angle_of(expr()));
        }
        expect(Tok::RParen, "`)`");
      }
      qubits(e.qubits);
      expect(Tok::Semi, "`;`");
      return g;
    }
    throw Error("unsupported gate `" + name + "`");
  }

  std::vector<Token> t_;
  size_t p_ = 0;
  bool has_qreg_ = false;
  std::string qreg_name_;
  uint32_t qreg_size_ = 0;
};

} // namespace

// ── Public API ─────────────────────────────────────────────────────────────

double QasmAngle::radians() const {
  return pi_multiple ? double(frac.num) / double(frac.den) * M_PI : number;
}

std::string QasmAngle::to_qasm() const {
  if (!pi_multiple)
    return shortest(number);
  const int64_t n = frac.num, d = frac.den;
  if (n == 0)
    return "0";
  const int64_t a = n < 0 ? -n : n;
  std::string s = (a == 1 ? std::string("pi") : std::to_string(a) + "*pi");
  if (d != 1)
    s += "/" + std::to_string(d);
  return n < 0 ? "-" + s : s;
}

std::string QasmGate::name() const {
  static const char *names[] = {"x", "y", "z", "h", "s", "t", "rx", "ry",
                                "rz", "u3", "cx", "cz", "cp", "swap", "ccx"};
  return names[int(op)];
}

std::string QasmGate::to_qasm() const {
  std::string s = name();
  if (!params.empty()) {
    s += "(";
    for (size_t i = 0; i < params.size(); ++i)
      s += (i ? "," : "") + params[i].to_qasm();
    s += ")";
  }
  for (size_t i = 0; i < qubits.size(); ++i)
    s += (i ? ",q[" : " q[") + std::to_string(qubits[i]) + "]";
  return s + ";";
}

uint32_t QasmCircuit::required_qreg_size() const {
  bool any = false;
  uint32_t hi = 0;
  for (const auto &g : gates)
    for (uint32_t q : g.qubits) {
      hi = any ? std::max(hi, q) : q;
      any = true;
    }
  return any ? hi + 1 : 0;
}

std::string QasmCircuit::serialize() const {
  std::string s = "OPENQASM 2.0;\nqreg q[" + std::to_string(required_qreg_size()) + "];\n";
  for (const auto &g : gates)
    s += g.to_qasm() + "\n";
  return s;
}

QasmCircuit parse_qasm(const std::string &text) { return Parser(lex(text)).circuit(); }

QuantumGate to_quantum_gate(const QasmGate &g) {
  const auto &q = g.qubits;
  auto a = [&](size_t i) { return g.params[i].radians(); };
  switch (g.op) {
  case QasmOp::X: return QuantumGate::x(q[0]);
  case QasmOp::Y: return QuantumGate::y(q[0]);
  case QasmOp::Z: return QuantumGate::z(q[0]);
  case QasmOp::H: return QuantumGate::h(q[0]);
  case QasmOp::S: return QuantumGate::s(q[0]);
  case QasmOp::T: return QuantumGate::t(q[0]);
  case QasmOp::RX: return QuantumGate::rx(a(0), q[0]);
  case QasmOp::RY: return QuantumGate::ry(a(0), q[0]);
  case QasmOp::RZ: return QuantumGate::rz(a(0), q[0]);
  case QasmOp::U3: return QuantumGate::u3(a(0), a(1), a(2), q[0]);
  case QasmOp::CX: return QuantumGate::cx(q[0], q[1]);
  case QasmOp::CZ: return QuantumGate::cz(q[0], q[1]);
  case QasmOp::CP: return QuantumGate::cp(a(0), q[0], q[1]);
  case QasmOp::SWAP: return QuantumGate::swap(q[0], q[1]);
  case QasmOp::CCX: return QuantumGate::ccx(q[0], q[1], q[2]);
  }
  throw Error("unknown gate kind");
}

QuantumCircuit circuit_from_qasm(const QasmCircuit &c) {
  QuantumCircuit out(c.required_qreg_size());
  for (const auto &g : c.gates)
    out.add(to_quantum_gate(g));
  return out;
}

CircuitGraph graph_from_qasm(const QasmCircuit &c) {
  CircuitGraph g;
  for (const auto &x : c.gates)
    g.insert_gate(std::make_shared<QuantumGate>(to_quantum_gate(x)));
  return g;
}

} // namespace cast
