// ini.h -- .ini reading with the semantics of the INIReader the reference links
// (jtilly/inih, a header-only C++ wrapper of benhoyt/inih; an un-vendored submodule of the
// reference, external/inih, version unpinned), and the reference's own Reader layer on top
// of it (SimInfo.h:100-263).
//
// Parser rules restated from inih's published behaviour (defaults of ini.h):
//  - the file is read in chunks of at most 199 characters per line (fgets into a 200-byte
//    buffer); a UTF-8 BOM on line 1 is skipped; every line is stripped of surrounding
//    whitespace;
//  - a line starting with ';' or '#' is a comment;
//  - "[section]" starts a section (an unterminated '[' is a parse error);
//  - "name = value" or "name : value" sets a value; an inline comment starts at a ';'
//    preceded by whitespace; name and value are stripped;
//  - an indented non-empty line after a value continues it (the handler is called again
//    with the same name);
//  - any other non-empty line is a parse error; parsing goes on, the first error line is
//    kept (the reference never checks it).
// INIReader stores every value under the lower-cased key "section=name"; a repeated key
// appends "\n" + value. Sections are remembered as written (case kept). Getters: Get
// (string, default if absent), GetInteger (strtol base 0 as long, default if no digits),
// GetFloat (strtof: single precision, default if no digits), GetReal (strtod), GetBoolean
// (true/yes/on/1, false/no/off/0, case-insensitive, else default).
#pragma once

#include <map>
#include <ostream>
#include <set>
#include <stdexcept>
#include <string>

namespace fv2dh {

class IniFile {
 public:
  IniFile() = default;
  explicit IniFile(const std::string &filename);
  // parse from text (same rules; used by tests)
  static IniFile from_string(const std::string &text);

  int parse_error() const { return error_; }  // 0, first bad line, or -1 (cannot open)
  const std::set<std::string> &sections() const { return sections_; }
  const std::map<std::string, std::string> &values() const { return values_; }

  std::string get(const std::string &section, const std::string &name, const std::string &def) const;
  long get_integer(const std::string &section, const std::string &name, long def) const;
  float get_float(const std::string &section, const std::string &name, float def) const;
  double get_real(const std::string &section, const std::string &name, double def) const;
  bool get_boolean(const std::string &section, const std::string &name, bool def) const;
  bool has_section(const std::string &section) const;
  bool has_value(const std::string &section, const std::string &name) const;

  static std::string make_key(const std::string &section, const std::string &name);

 private:
  void parse_lines(const std::string &text);
  void on_value(const std::string &section, const std::string &name, const std::string &value);
  int error_ = 0;
  std::map<std::string, std::string> values_;
  std::set<std::string> sections_;
};

// The reference's Reader (SimInfo.h:122-263): reads through IniFile, remembers every
// parameter it was asked for (lower-cased section and name, the value used formatted as
// text, whether it came from the file), refuses to register one twice, maps enum values
// and prints the parameters in .ini form.
class Reader {
 public:
  struct Value {
    std::string value;
    bool from_file = false;
    bool is_default_value = true;
  };

  Reader() = default;
  explicit Reader(const std::string &filename) : ini(filename) {}
  explicit Reader(IniFile f) : ini(std::move(f)) {}

  bool get_boolean(const std::string &section, const std::string &name, bool def);
  int get_integer(const std::string &section, const std::string &name, int def);
  double get_float(const std::string &section, const std::string &name, double def);  // through float, as GetFloat
  std::string get(const std::string &section, const std::string &name, const std::string &def);

  template <typename T>
  T get_map_value(const std::map<std::string, T> &map, const std::string &section, const std::string &name,
                  const std::string &def) {
    const std::string v = get(section, name, def);
    auto it = map.find(v);
    if (it == map.end()) {
      std::string allowed = "\nallowed values: ";
      for (const auto &e : map) allowed += e.first + ", ";
      throw std::runtime_error("bad parameter for " + name + ": " + allowed);
    }
    return it->second;
  }

  // "; Parameters used for the problem: <problem>" then every read section present in the
  // file, name = value per line, values not from the file marked "; default".
  void output_values(std::ostream &o) const;

  IniFile ini;
  std::map<std::string, std::map<std::string, Value>> registered;

 private:
  void register_value(std::string section, std::string name, const std::string &text, bool is_default);
};

std::string lower(std::string s);
std::string format_float(double v);  // std::scientific, 12 digits, as the reference registers floats

}  // namespace fv2dh
