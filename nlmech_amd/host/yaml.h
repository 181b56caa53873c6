// YAML subset reader for NLMech input decks.
//
// The decks (examples/*/input.yaml, tools/*/input*.yaml) use block mappings,
// block sequences ("- item"), flow sequences ("[a, b]"), plain or quoted
// scalars and "#" comments.  This reader covers exactly that subset; it is
// what inp::Input reads through yaml-cpp in the reference
// (src/inp/input.cpp:124-899).  A key given twice in one mapping keeps its
// first value, as yaml-cpp's lookup does (examples/fdModel/restart/input.yaml
// repeats "Parameters").
#pragma once

#include <cstddef>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace nlm {
namespace yaml {

class Error : public std::runtime_error {
 public:
  using std::runtime_error::runtime_error;
};

class Node {
 public:
  enum Kind { Null, Scalar, Sequence, Map };

  Node() = default;

  Kind kind() const { return kind_; }
  bool is_null() const { return kind_ == Null; }
  bool is_scalar() const { return kind_ == Scalar; }
  bool is_sequence() const { return kind_ == Sequence; }
  bool is_map() const { return kind_ == Map; }
  // True when the key exists, also with an empty value (yaml-cpp's
  // IsDefined, which `if (config["Restart"])` tests in the reference).
  explicit operator bool() const { return defined_; }

  // Mapping lookup; a missing key (or a non-mapping node) gives a null node.
  const Node& operator[](const std::string& key) const;
  // Sequence element; out of range gives a null node.
  const Node& operator[](size_t i) const;
  // Number of sequence elements or mapping entries (0 for scalars).
  size_t size() const;
  const std::vector<Node>& items() const { return seq_; }
  const std::vector<std::pair<std::string, Node>>& entries() const {
    return map_;
  }
  int line() const { return line_; }

  // Scalar conversions; throw Error with the line number on a mismatch.
  const std::string& str() const;
  double as_double() const;
  long long as_int() const;
  size_t as_size() const;
  bool as_bool() const;
  // A sequence of numbers (a scalar gives a one-element vector).
  std::vector<double> as_doubles() const;
  std::vector<std::string> as_strings() const;

  // Builders used by the parser.
  static Node empty(int line);
  static Node scalar(std::string s, int line);
  static Node sequence(int line);
  static Node mapping(int line);
  void push(Node n) { seq_.push_back(std::move(n)); }
  void set(std::string key, Node n);

 private:
  Kind kind_ = Null;
  bool defined_ = false;
  std::string scalar_;
  std::vector<Node> seq_;
  std::vector<std::pair<std::string, Node>> map_;
  int line_ = 0;
};

Node load(const std::string& text);
Node load_file(const std::string& path);

}  // namespace yaml
}  // namespace nlm
