// ini.cpp - Settings.ini reader; behaviour follows F/INIlib/ini.c:64-170 and
// F/INIlib/INIReader.cpp:14-85 (see ini.h for the list of rules).
#include "ini.h"

#include <cctype>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <vector>

namespace fdtd2d {
namespace {

constexpr size_t kMaxLine = 200;  // INI_MAX_LINE (F/INIlib/ini.h:76)
constexpr size_t kMaxSection = 50;
constexpr size_t kMaxName = 50;

bool is_space(char c) { return std::isspace(static_cast<unsigned char>(c)) != 0; }

std::string rstrip(const std::string &s) {
  size_t e = s.size();
  while (e > 0 && is_space(s[e - 1])) --e;
  return s.substr(0, e);
}

size_t lskip(const std::string &s, size_t p) {
  while (p < s.size() && is_space(s[p])) ++p;
  return p;
}

// First position of c, or of a ';' that follows whitespace, from p on.
size_t find_char_or_comment(const std::string &s, size_t p, char c) {
  bool was_ws = false;
  while (p < s.size() && s[p] != c && !(was_ws && s[p] == ';')) {
    was_ws = is_space(s[p]);
    ++p;
  }
  return p;
}

// fgets(buf, kMaxLine, f) as a splitter: each chunk ends after '\n' or after
// kMaxLine-1 characters.
std::vector<std::string> fgets_chunks(const std::string &text) {
  std::vector<std::string> out;
  size_t p = 0;
  while (p < text.size()) {
    size_t n = 0;
    while (p + n < text.size() && n < kMaxLine - 1) {
      char c = text[p + n];
      ++n;
      if (c == '\n') break;
    }
    out.push_back(text.substr(p, n));
    p += n;
  }
  return out;
}

std::string truncate0(const std::string &s, size_t size) {
  return s.size() >= size ? s.substr(0, size - 1) : s;
}

}  // namespace

IniFile::IniFile(const std::string &path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) {
    status_ = -1;
    return;
  }
  std::stringstream ss;
  ss << f.rdbuf();
  parse_text(ss.str());
}

IniFile IniFile::from_text(const std::string &text) {
  IniFile r;
  r.parse_text(text);
  return r;
}

std::string IniFile::key(const std::string &section, const std::string &name) {
  std::string k = section + "=" + name;
  for (auto &c : k) c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
  return k;
}

void IniFile::parse_text(const std::string &text) {
  std::string section, prev_name;
  int lineno = 0, error = 0;
  for (const std::string &raw : fgets_chunks(text)) {
    ++lineno;
    std::string line = raw;
    size_t start0 = 0;
    if (lineno == 1 && line.size() >= 3 && static_cast<unsigned char>(line[0]) == 0xEF &&
        static_cast<unsigned char>(line[1]) == 0xBB && static_cast<unsigned char>(line[2]) == 0xBF)
      start0 = 3;
    line = rstrip(line);
    size_t start = lskip(line, start0);
    const bool empty = start >= line.size();
    const char c0 = empty ? '\0' : line[start];
    if (c0 == ';' || c0 == '#') {
      // comment line
    } else if (!prev_name.empty() && !empty && start > 0) {
      // continuation of the previous name's value
      std::string v = line.substr(start);
      std::string k = key(section, prev_name);
      if (!values_[k].empty()) values_[k] += "\n";
      values_[k] += v;
    } else if (c0 == '[') {
      size_t end = find_char_or_comment(line, start + 1, ']');
      if (end < line.size() && line[end] == ']') {
        section = truncate0(line.substr(start + 1, end - start - 1), kMaxSection);
        prev_name.clear();
      } else if (!error) {
        error = lineno;
      }
    } else if (!empty) {
      size_t end = find_char_or_comment(line, start, '=');
      if (!(end < line.size() && line[end] == '=')) end = find_char_or_comment(line, start, ':');
      if (end < line.size() && (line[end] == '=' || line[end] == ':')) {
        std::string name = rstrip(line.substr(start, end - start));
        size_t vs = lskip(line, end + 1);
        size_t ve = find_char_or_comment(line, vs, '\0');
        std::string value = rstrip(line.substr(vs, ve - vs));
        prev_name = truncate0(name, kMaxName);
        std::string k = key(section, name);
        if (!values_[k].empty()) values_[k] += "\n";
        values_[k] += value;
      } else if (!error) {
        error = lineno;
      }
    }
  }
  status_ = error;
}

bool IniFile::has(const std::string &section, const std::string &name) const {
  return values_.count(key(section, name)) != 0;
}

std::string IniFile::get(const std::string &section, const std::string &name,
                         const std::string &def) const {
  auto it = values_.find(key(section, name));
  return it == values_.end() ? def : it->second;
}

long IniFile::get_int(const std::string &section, const std::string &name, long def) const {
  std::string v = get(section, name, "");
  const char *s = v.c_str();
  char *end = nullptr;
  long n = std::strtol(s, &end, 0);
  return end > s ? n : def;
}

double IniFile::get_real(const std::string &section, const std::string &name, double def) const {
  std::string v = get(section, name, "");
  const char *s = v.c_str();
  char *end = nullptr;
  double n = std::strtod(s, &end);
  return end > s ? n : def;
}

bool IniFile::get_bool(const std::string &section, const std::string &name, bool def) const {
  std::string v = get(section, name, "");
  for (auto &c : v) c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
  if (v == "true" || v == "yes" || v == "on" || v == "1") return true;
  if (v == "false" || v == "no" || v == "off" || v == "0") return false;
  return def;
}

}  // namespace fdtd2d
