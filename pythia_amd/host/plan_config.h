// Plan configuration in libconfig syntax, with the libconfig::Config /
// libconfig::Setting interface that Pythia's operators call
// (operators/*.cpp: node["key"], exists, lookupValue, getLength, add,
// remove, implicit scalar conversions).  Implemented from the published
// file format: settings `name = value;` or `name : value;`, groups { },
// lists ( ), arrays [ ], ints (decimal / hex, `L` suffix for 64-bit),
// floats, booleans, strings with escapes and adjacent-literal
// concatenation, comments #, //, /* */.
#pragma once

#include <cstdint>
#include <cstdio>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace libconfig {

class ConfigException : public std::runtime_error {
  public:
    explicit ConfigException(const std::string& m) : std::runtime_error(m) {}
};
class SettingException : public ConfigException {
  public:
    explicit SettingException(const std::string& m) : ConfigException(m) {}
};
class SettingNotFoundException : public SettingException {
  public:
    explicit SettingNotFoundException(const std::string& path)
        : SettingException("setting not found: " + path) {}
};
class SettingTypeException : public SettingException {
  public:
    explicit SettingTypeException(const std::string& path) : SettingException("setting type mismatch: " + path) {}
};
class SettingNameException : public SettingException {
  public:
    explicit SettingNameException(const std::string& n) : SettingException("bad or duplicate setting name: " + n) {}
};
class ParseException : public ConfigException {
  public:
    ParseException(int line, const std::string& what)
        : ConfigException("parse error at line " + std::to_string(line) + ": " + what), line_(line) {}
    int getLine() const { return line_; }

  private:
    int line_;
};

class Config;

class Setting {
  public:
    enum Type { TypeNone = 0, TypeGroup, TypeInt, TypeInt64, TypeFloat, TypeString, TypeBoolean, TypeArray, TypeList };

    Type getType() const { return type_; }
    const char* getName() const { return name_.empty() ? nullptr : name_.c_str(); }
    std::string getPath() const;
    bool isGroup() const { return type_ == TypeGroup; }
    bool isArray() const { return type_ == TypeArray; }
    bool isList() const { return type_ == TypeList; }
    bool isAggregate() const { return type_ == TypeGroup || type_ == TypeArray || type_ == TypeList; }
    bool isScalar() const { return !isAggregate() && type_ != TypeNone; }
    bool isNumber() const { return type_ == TypeInt || type_ == TypeInt64 || type_ == TypeFloat; }
    int getLength() const { return isAggregate() ? (int)children_.size() : 0; }
    bool isRoot() const { return parent_ == nullptr; }
    Setting& getParent() const;

    Setting& operator[](int idx) const;
    Setting& operator[](const char* name) const;
    Setting& operator[](const std::string& name) const { return (*this)[name.c_str()]; }
    bool exists(const char* name) const;
    bool exists(const std::string& name) const { return exists(name.c_str()); }

    // Scalar reads (libconfig conversion rules: any integer width, int <-> float).
    operator bool() const;
    operator int() const;
    operator unsigned int() const;
    operator long() const;
    operator unsigned long() const;
    operator long long() const;
    operator unsigned long long() const;
    operator double() const;
    operator float() const;
    operator std::string() const;
    operator const char*() const;

    template <typename T>
    bool lookupValue(const char* name, T& value) const {
        if (!isGroup() || !exists(name)) return false;
        const Setting& s = (*this)[name];
        if (!s.isScalar()) return false;
        try {
            value = (T)s;
        } catch (const SettingTypeException&) {
            return false;
        }
        return true;
    }
    bool lookupValue(const char* name, std::string& value) const;

    Setting& operator=(bool v);
    Setting& operator=(int v);
    Setting& operator=(long v);
    Setting& operator=(long long v);
    Setting& operator=(unsigned int v) { return *this = (long long)v; }
    Setting& operator=(double v);
    Setting& operator=(const char* v);
    Setting& operator=(const std::string& v) { return *this = v.c_str(); }

    // Adds a named child to a group, or an unnamed element to a list/array.
    Setting& add(const char* name, Type t);
    Setting& add(const std::string& name, Type t) { return add(name.c_str(), t); }
    Setting& add(Type t);
    void remove(const char* name);
    void remove(unsigned int idx);

    int64_t intValue() const { return ival_; }
    void write(std::string& out, int indent) const;

    Setting(const Setting&) = delete;
    Setting& operator=(const Setting&) = delete;

  private:
    friend class Config;
    friend class Parser;
    Setting(Setting* parent, std::string name, Type t) : parent_(parent), name_(std::move(name)), type_(t) {}
    int64_t as_int64() const;
    void set_scalar_type(Type t);

    Setting* parent_;
    std::string name_;
    Type type_;
    int64_t ival_ = 0;
    double fval_ = 0;
    bool bval_ = false;
    std::string sval_;
    std::vector<std::unique_ptr<Setting>> children_;
};

class Config {
  public:
    Config();
    ~Config();
    void readFile(const char* filename);
    void readString(const char* text);
    void readString(const std::string& text) { readString(text.c_str()); }
    void writeFile(const char* filename) const;
    std::string toString() const;
    Setting& getRoot() const { return *root_; }
    Setting& lookup(const char* path) const;
    Setting& lookup(const std::string& path) const { return lookup(path.c_str()); }
    bool exists(const char* path) const;
    template <typename T>
    bool lookupValue(const char* path, T& v) const {
        if (!exists(path)) return false;
        try {
            v = (T)lookup(path);
        } catch (const SettingTypeException&) {
            return false;
        }
        return true;
    }

    Config(const Config&) = delete;
    Config& operator=(const Config&) = delete;

  private:
    std::unique_ptr<Setting> root_;
};

}  // namespace libconfig
