// YAML subset reader for NLMech input decks (see yaml.h for the subset).
#include "yaml.h"

#include <cerrno>
#include <cstdlib>
#include <fstream>
#include <sstream>

namespace nlm {
namespace yaml {

namespace {

const Node& null_node() {
  static const Node n;
  return n;
}

std::string trim(const std::string& s) {
  size_t a = 0, b = s.size();
  while (a < b && (s[a] == ' ' || s[a] == '\t' || s[a] == '\r')) a++;
  while (b > a && (s[b - 1] == ' ' || s[b - 1] == '\t' || s[b - 1] == '\r'))
    b--;
  return s.substr(a, b - a);
}

[[noreturn]] void fail(int line, const std::string& what) {
  throw Error("YAML line " + std::to_string(line) + ": " + what);
}

struct Line {
  int indent;
  std::string text;  // without indentation, comment and trailing blanks
  int no;
};

// Removes a "#" comment that starts the line or follows a blank, outside
// quotes.
std::string strip_comment(const std::string& s) {
  char quote = 0;
  for (size_t i = 0; i < s.size(); i++) {
    char c = s[i];
    if (quote) {
      if (c == quote) quote = 0;
      continue;
    }
    if (c == '"' || c == '\'') {
      quote = c;
    } else if (c == '#' && (i == 0 || s[i - 1] == ' ' || s[i - 1] == '\t')) {
      return s.substr(0, i);
    }
  }
  return s;
}

int bracket_depth(const std::string& s) {
  int d = 0;
  char quote = 0;
  for (char c : s) {
    if (quote) {
      if (c == quote) quote = 0;
      continue;
    }
    if (c == '"' || c == '\'') quote = c;
    else if (c == '[' || c == '{') d++;
    else if (c == ']' || c == '}') d--;
  }
  return d;
}

std::vector<Line> split_lines(const std::string& text) {
  std::vector<Line> out;
  std::istringstream is(text);
  std::string raw;
  int no = 0;
  while (std::getline(is, raw)) {
    no++;
    std::string s = strip_comment(raw);
    size_t ind = 0;
    while (ind < s.size() && s[ind] == ' ') ind++;
    if (ind < s.size() && s[ind] == '\t')
      fail(no, "tab in indentation");
    std::string t = trim(s.substr(ind));
    if (t.empty() || t == "---" || t == "...") continue;
    // a flow collection left open continues on the next lines
    if (!out.empty() && bracket_depth(out.back().text) > 0) {
      out.back().text += " " + t;
      continue;
    }
    out.push_back({int(ind), t, no});
  }
  return out;
}

// Position of the ':' that ends a mapping key (followed by a blank or the
// end of the text, outside quotes and brackets); npos if none.
size_t key_colon(const std::string& t) {
  char quote = 0;
  int depth = 0;
  for (size_t i = 0; i < t.size(); i++) {
    char c = t[i];
    if (quote) {
      if (c == quote) quote = 0;
      continue;
    }
    if (c == '"' || c == '\'') quote = c;
    else if (c == '[' || c == '{') depth++;
    else if (c == ']' || c == '}') depth--;
    else if (c == ':' && depth == 0 &&
             (i + 1 == t.size() || t[i + 1] == ' '))
      return i;
  }
  return std::string::npos;
}

std::string unquote(const std::string& s, int line) {
  if (s.size() >= 2 && (s[0] == '"' || s[0] == '\'')) {
    if (s.back() != s[0]) fail(line, "unterminated quoted scalar");
    std::string body = s.substr(1, s.size() - 2);
    if (s[0] == '\'') {
      std::string out;
      for (size_t i = 0; i < body.size(); i++) {
        out += body[i];
        if (body[i] == '\'' && i + 1 < body.size() && body[i + 1] == '\'') i++;
      }
      return out;
    }
    std::string out;
    for (size_t i = 0; i < body.size(); i++) {
      if (body[i] == '\\' && i + 1 < body.size()) {
        char e = body[++i];
        out += e == 'n' ? '\n' : e == 't' ? '\t' : e;
      } else {
        out += body[i];
      }
    }
    return out;
  }
  return s;
}

Node parse_inline(const std::string& t, int line);

// Splits the inside of a flow collection at top-level commas.
std::vector<std::string> split_flow(const std::string& body, int line) {
  std::vector<std::string> parts;
  std::string cur;
  int depth = 0;
  char quote = 0;
  for (char c : body) {
    if (quote) {
      cur += c;
      if (c == quote) quote = 0;
      continue;
    }
    if (c == '"' || c == '\'') quote = c;
    if (c == '[' || c == '{') depth++;
    if (c == ']' || c == '}') depth--;
    if (c == ',' && depth == 0) {
      parts.push_back(trim(cur));
      cur.clear();
      continue;
    }
    cur += c;
  }
  if (depth != 0 || quote) fail(line, "unbalanced flow collection");
  std::string last = trim(cur);
  if (!last.empty()) parts.push_back(last);
  return parts;
}

Node parse_inline(const std::string& t, int line) {
  if (t.empty() || t == "~" || t == "null") return Node::empty(line);
  if (t[0] == '[') {
    if (t.back() != ']') fail(line, "flow sequence not closed");
    Node seq = Node::sequence(line);
    for (const auto& p : split_flow(t.substr(1, t.size() - 2), line))
      seq.push(parse_inline(p, line));
    return seq;
  }
  if (t[0] == '{') {
    if (t.back() != '}') fail(line, "flow mapping not closed");
    Node map = Node::mapping(line);
    for (const auto& p : split_flow(t.substr(1, t.size() - 2), line)) {
      size_t c = key_colon(p);
      if (c == std::string::npos) fail(line, "flow mapping entry without ':'");
      map.set(unquote(trim(p.substr(0, c)), line),
              parse_inline(trim(p.substr(c + 1)), line));
    }
    return map;
  }
  return Node::scalar(unquote(t, line), line);
}

class Parser {
 public:
  explicit Parser(std::vector<Line> lines) : l_(std::move(lines)) {}

  Node document() {
    if (l_.empty()) return Node::mapping(0);
    Node n = block(l_[0].indent);
    if (pos_ < l_.size()) fail(l_[pos_].no, "unexpected indentation");
    return n;
  }

 private:
  std::vector<Line> l_;
  size_t pos_ = 0;

  static bool is_item(const std::string& t) {
    return t == "-" || (t.size() >= 2 && t[0] == '-' && t[1] == ' ');
  }

  Node block(int indent) {
    if (is_item(l_[pos_].text)) return sequence(indent);
    return mapping(indent);
  }

  Node sequence(int indent) {
    Node seq = Node::sequence(l_[pos_].no);
    while (pos_ < l_.size() && l_[pos_].indent == indent &&
           is_item(l_[pos_].text)) {
      Line& cur = l_[pos_];
      std::string rest = trim(cur.text.substr(1));
      if (rest.empty()) {
        pos_++;
        if (pos_ < l_.size() && l_[pos_].indent > indent)
          seq.push(block(l_[pos_].indent));
        else
          seq.push(Node::empty(cur.no));
      } else if (key_colon(rest) != std::string::npos && rest[0] != '[' &&
                 rest[0] != '{' && rest[0] != '"' && rest[0] != '\'') {
        // "- key: value" opens a mapping indented past the dash
        int ind = indent + int(cur.text.size() - rest.size());
        cur.indent = ind;
        cur.text = rest;
        seq.push(mapping(ind));
      } else {
        seq.push(parse_inline(rest, cur.no));
        pos_++;
      }
    }
    return seq;
  }

  Node mapping(int indent) {
    Node map = Node::mapping(l_[pos_].no);
    while (pos_ < l_.size() && l_[pos_].indent == indent) {
      const Line cur = l_[pos_];
      if (is_item(cur.text)) fail(cur.no, "sequence item inside a mapping");
      size_t c = key_colon(cur.text);
      if (c == std::string::npos)
        fail(cur.no, "expected 'key: value', got '" + cur.text + "'");
      std::string key = unquote(trim(cur.text.substr(0, c)), cur.no);
      std::string rest = trim(cur.text.substr(c + 1));
      pos_++;
      if (!rest.empty()) {
        map.set(key, parse_inline(rest, cur.no));
        continue;
      }
      if (pos_ < l_.size() && l_[pos_].indent > indent) {
        map.set(key, block(l_[pos_].indent));
      } else if (pos_ < l_.size() && l_[pos_].indent == indent &&
                 is_item(l_[pos_].text)) {
        map.set(key, sequence(indent));
      } else {
        map.set(key, Node::empty(cur.no));
      }
    }
    if (pos_ < l_.size() && l_[pos_].indent > indent)
      fail(l_[pos_].no, "unexpected indentation");
    return map;
  }
};

}  // namespace

Node Node::empty(int line) {
  Node n;
  n.defined_ = true;
  n.line_ = line;
  return n;
}

Node Node::scalar(std::string s, int line) {
  Node n = empty(line);
  n.kind_ = Scalar;
  n.scalar_ = std::move(s);
  return n;
}

Node Node::sequence(int line) {
  Node n = empty(line);
  n.kind_ = Sequence;
  return n;
}

Node Node::mapping(int line) {
  Node n = empty(line);
  n.kind_ = Map;
  return n;
}

void Node::set(std::string key, Node n) {
  for (const auto& e : map_)
    if (e.first == key) return;  // first occurrence wins
  map_.emplace_back(std::move(key), std::move(n));
}

const Node& Node::operator[](const std::string& key) const {
  if (kind_ != Map) return null_node();
  for (const auto& e : map_)
    if (e.first == key) return e.second;
  return null_node();
}

const Node& Node::operator[](size_t i) const {
  if (kind_ != Sequence || i >= seq_.size()) return null_node();
  return seq_[i];
}

size_t Node::size() const {
  return kind_ == Sequence ? seq_.size() : kind_ == Map ? map_.size() : 0;
}

const std::string& Node::str() const {
  if (kind_ != Scalar) fail(line_, "expected a scalar");
  return scalar_;
}

double Node::as_double() const {
  const std::string& s = str();
  errno = 0;
  char* end = nullptr;
  double v = std::strtod(s.c_str(), &end);
  if (end == s.c_str() || *end != '\0' || errno == ERANGE)
    fail(line_, "'" + s + "' is not a number");
  return v;
}

long long Node::as_int() const {
  const std::string& s = str();
  errno = 0;
  char* end = nullptr;
  long long v = std::strtoll(s.c_str(), &end, 10);
  if (end == s.c_str() || *end != '\0' || errno == ERANGE)
    fail(line_, "'" + s + "' is not an integer");
  return v;
}

size_t Node::as_size() const {
  long long v = as_int();
  if (v < 0) fail(line_, "'" + str() + "' is negative");
  return size_t(v);
}

bool Node::as_bool() const {
  const std::string& s = str();
  static const char* yes[] = {"true", "True", "TRUE", "yes", "Yes",
                              "YES",  "on",   "On",   "ON",  "y", "Y"};
  static const char* no[] = {"false", "False", "FALSE", "no", "No",
                             "NO",    "off",   "Off",   "OFF", "n", "N"};
  for (const char* t : yes)
    if (s == t) return true;
  for (const char* t : no)
    if (s == t) return false;
  fail(line_, "'" + s + "' is not a boolean");
}

std::vector<double> Node::as_doubles() const {
  std::vector<double> out;
  if (kind_ == Scalar) {
    out.push_back(as_double());
  } else if (kind_ == Sequence) {
    for (const auto& n : seq_) out.push_back(n.as_double());
  } else if (kind_ == Map) {
    fail(line_, "expected a sequence of numbers");
  }
  return out;
}

std::vector<std::string> Node::as_strings() const {
  std::vector<std::string> out;
  if (kind_ == Scalar) {
    out.push_back(scalar_);
  } else if (kind_ == Sequence) {
    for (const auto& n : seq_) out.push_back(n.str());
  } else if (kind_ == Map) {
    fail(line_, "expected a sequence of strings");
  }
  return out;
}

Node load(const std::string& text) {
  return Parser(split_lines(text)).document();
}

Node load_file(const std::string& path) {
  std::ifstream f(path);
  if (!f) throw Error("cannot open " + path);
  std::stringstream ss;
  ss << f.rdbuf();
  return load(ss.str());
}

}  // namespace yaml
}  // namespace nlm
