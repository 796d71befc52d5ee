// ini.cpp -- inih-compatible parser and the reference's Reader layer (see ini.h).
#include "ini.h"

#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <sstream>

namespace fv2dh {

namespace {

constexpr size_t kMaxLine = 200;  // inih INI_MAX_LINE: fgets buffer, 199 characters per read
constexpr size_t kMaxSection = 50;
constexpr size_t kMaxName = 50;

std::string rstrip(const std::string &s) {
  size_t e = s.size();
  while (e > 0 && std::isspace((unsigned char)s[e - 1])) --e;
  return s.substr(0, e);
}

size_t lskip(const std::string &s, size_t i = 0) {
  while (i < s.size() && std::isspace((unsigned char)s[i])) ++i;
  return i;
}

// First position in s[from..] of a char in `chars` (if given), or of an inline comment
// (';' preceded by whitespace); s.size() if none.
size_t find_chars_or_comment(const std::string &s, size_t from, const char *chars) {
  bool was_space = false;
  size_t i = from;
  while (i < s.size() && (!chars || !std::strchr(chars, s[i])) && !(was_space && s[i] == ';')) {
    was_space = std::isspace((unsigned char)s[i]) != 0;
    ++i;
  }
  return i;
}

}  // namespace

std::string lower(std::string s) {
  std::transform(s.begin(), s.end(), s.begin(), [](unsigned char c) { return (char)std::tolower(c); });
  return s;
}

std::string format_float(double v) {
  std::ostringstream os;
  os << std::scientific << std::setprecision(12) << v;
  return os.str();
}

IniFile::IniFile(const std::string &filename) {
  std::ifstream f(filename, std::ios::binary);
  if (!f) {
    error_ = -1;
    return;
  }
  std::ostringstream ss;
  ss << f.rdbuf();
  parse_lines(ss.str());
}

IniFile IniFile::from_string(const std::string &text) {
  IniFile f;
  f.parse_lines(text);
  return f;
}

std::string IniFile::make_key(const std::string &section, const std::string &name) {
  return lower(section + "=" + name);
}

void IniFile::on_value(const std::string &section, const std::string &name, const std::string &value) {
  std::string &v = values_[make_key(section, name)];
  if (!v.empty()) v += "\n";
  v += value;
  sections_.insert(section);
}

void IniFile::parse_lines(const std::string &text) {
  std::string section, prev_name;
  int lineno = 0;
  size_t pos = 0;
  while (pos < text.size()) {
    // one fgets() read: up to and including '\n', at most kMaxLine - 1 characters
    size_t end = pos;
    while (end < text.size() && end - pos < kMaxLine - 1) {
      if (text[end++] == '\n') break;
    }
    std::string line = text.substr(pos, end - pos);
    pos = end;
    ++lineno;
    size_t start = 0;
    if (lineno == 1 && line.size() >= 3 && (unsigned char)line[0] == 0xEF && (unsigned char)line[1] == 0xBB &&
        (unsigned char)line[2] == 0xBF)
      start = 3;
    line = rstrip(line);
    start = lskip(line, start);
    const bool indented = start > 0;  // on line 1 (BOM) no value precedes, so no continuation
    if (start >= line.size()) continue;  // empty
    const char c = line[start];
    if (c == ';' || c == '#') continue;  // comment line
    if (!prev_name.empty() && indented) {
      // continuation of the previous value
      size_t e = find_chars_or_comment(line, start, nullptr);
      on_value(section, prev_name, rstrip(line.substr(start, e - start)));
      continue;
    }
    if (c == '[') {
      size_t e = find_chars_or_comment(line, start + 1, "]");
      if (e < line.size() && line[e] == ']') {
        section = line.substr(start + 1, e - start - 1).substr(0, kMaxSection - 1);
        prev_name.clear();
      } else if (!error_) {
        error_ = lineno;  // no ']'
      }
      continue;
    }
    size_t e = find_chars_or_comment(line, start, "=:");
    if (e < line.size() && (line[e] == '=' || line[e] == ':')) {
      std::string name = rstrip(line.substr(start, e - start));
      size_t vs = lskip(line, e + 1);
      size_t ve = find_chars_or_comment(line, vs, nullptr);
      std::string value = rstrip(line.substr(vs, ve - vs));
      prev_name = name.substr(0, kMaxName - 1);
      on_value(section, name, value);
    } else if (!error_) {
      error_ = lineno;  // no '=' or ':'
    }
  }
}

std::string IniFile::get(const std::string &section, const std::string &name, const std::string &def) const {
  auto it = values_.find(make_key(section, name));
  return it == values_.end() ? def : it->second;
}

long IniFile::get_integer(const std::string &section, const std::string &name, long def) const {
  const std::string v = get(section, name, "");
  const char *s = v.c_str();
  char *end = nullptr;
  const long n = std::strtol(s, &end, 0);
  return end > s ? n : def;
}

float IniFile::get_float(const std::string &section, const std::string &name, float def) const {
  const std::string v = get(section, name, "");
  const char *s = v.c_str();
  char *end = nullptr;
  const float n = std::strtof(s, &end);
  return end > s ? n : def;
}

double IniFile::get_real(const std::string &section, const std::string &name, double def) const {
  const std::string v = get(section, name, "");
  const char *s = v.c_str();
  char *end = nullptr;
  const double n = std::strtod(s, &end);
  return end > s ? n : def;
}

bool IniFile::get_boolean(const std::string &section, const std::string &name, bool def) const {
  const std::string v = lower(get(section, name, ""));
  if (v == "true" || v == "yes" || v == "on" || v == "1") return true;
  if (v == "false" || v == "no" || v == "off" || v == "0") return false;
  return def;
}

bool IniFile::has_section(const std::string &section) const {
  const std::string key = make_key(section, "");
  auto it = values_.lower_bound(key);
  if (it == values_.end()) return false;
  return it->first.compare(0, key.size(), key) == 0;
}

bool IniFile::has_value(const std::string &section, const std::string &name) const {
  return values_.count(make_key(section, name)) != 0;
}

// ------------------------------------------------------------------ Reader

void Reader::register_value(std::string section, std::string name, const std::string &text, bool is_default) {
  section = lower(section);
  name = lower(name);
  auto s = registered.find(section);
  if (s != registered.end() && s->second.count(name))
    throw std::runtime_error(std::string("parameter already set : ") + name);
  Value &v = registered[section][name];
  if (ini.has_section(section) && ini.has_value(section, name)) {
    v.from_file = true;
    v.is_default_value = is_default;
  }
  v.value = text;
}

bool Reader::get_boolean(const std::string &section, const std::string &name, bool def) {
  const bool r = ini.get_boolean(section, name, def);
  register_value(section, name, r ? "true" : "false", r == def);
  return r;
}

int Reader::get_integer(const std::string &section, const std::string &name, int def) {
  const int r = (int)ini.get_integer(section, name, def);
  register_value(section, name, std::to_string(r), r == def);
  return r;
}

double Reader::get_float(const std::string &section, const std::string &name, double def) {
  const double r = (double)ini.get_float(section, name, (float)def);
  register_value(section, name, format_float(r), r == def);
  return r;
}

std::string Reader::get(const std::string &section, const std::string &name, const std::string &def) {
  std::string r = ini.get(section, name, def);
  register_value(section, name, r, r == def);
  return r;
}

void Reader::output_values(std::ostream &o) const {
  const std::string::size_type name_w = 26, value_w = 20;
  std::string problem;
  auto ps = registered.find("physics");
  if (ps != registered.end()) {
    auto pp = ps->second.find("problem");
    if (pp != ps->second.end()) problem = pp->second.value;
  }
  const auto flags = o.flags();
  o << "; Parameters used for the problem: " << problem << std::endl;
  o << std::left;
  for (const auto &sec : registered) {
    if (!ini.has_section(sec.first)) continue;
    o << "\n[" << sec.first << "]" << std::endl;
    for (const auto &var : sec.second) {
      o << std::setw((int)std::max(var.first.size(), name_w)) << var.first << " = "
        << std::setw((int)std::max(var.second.value.size(), value_w)) << var.second.value
        << (var.second.from_file ? "" : " ; default ") << std::endl;
    }
  }
  o.flags(flags);
}

}  // namespace fv2dh
