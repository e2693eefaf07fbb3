// ini.h - Settings.ini reader with the semantics of the reference's vendored
// inih + INIReader (F/INIlib/ini.c, F/INIlib/INIReader.cpp):
//  * lines are read 199 characters at a time (INI_MAX_LINE 200): the rest of
//    a longer line is parsed as a line of its own;
//  * a UTF-8 BOM on line 1 is skipped; ';' or '#' at line start is a comment;
//  * an inline ';' is a comment only when preceded by whitespace;
//  * "name = value" or "name : value"; leading-whitespace lines continue the
//    previous name (the values are joined with '\n');
//  * section and key lookups are case-insensitive; duplicate keys join '\n';
//  * GetInteger is strtol(..., 0) (hex, octal, "0.5" -> 0), GetReal strtod,
//    GetBoolean true/yes/on/1 and false/no/off/0; anything unparsable falls
//    back to the default; parse errors record the first bad line and go on.
#ifndef FDTD2D_INI_H
#define FDTD2D_INI_H

#include <map>
#include <string>

namespace fdtd2d {

class IniFile {
 public:
  // Parse a file; status(): 0 ok, -1 cannot open, >0 first line with an error.
  explicit IniFile(const std::string &path);
  // Parse text already in memory (same rules).
  static IniFile from_text(const std::string &text);
  int status() const { return status_; }
  bool has(const std::string &section, const std::string &name) const;
  std::string get(const std::string &section, const std::string &name,
                  const std::string &def) const;
  long get_int(const std::string &section, const std::string &name, long def) const;
  double get_real(const std::string &section, const std::string &name, double def) const;
  bool get_bool(const std::string &section, const std::string &name, bool def) const;

 private:
  IniFile() = default;
  void parse_text(const std::string &text);
  static std::string key(const std::string &section, const std::string &name);
  std::map<std::string, std::string> values_;
  int status_ = 0;
};

}  // namespace fdtd2d
#endif
