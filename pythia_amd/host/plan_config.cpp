// Recursive-descent reader/writer for libconfig-syntax plan files.
#include "plan_config.h"

#include <cctype>
#include <cerrno>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

namespace libconfig {

// ------------------------------------------------------------ Setting
std::string Setting::getPath() const {
    if (!parent_) return "";
    std::string p = parent_->getPath();
    std::string me = name_;
    if (me.empty()) {
        for (size_t i = 0; i < parent_->children_.size(); ++i)
            if (parent_->children_[i].get() == this) me = "[" + std::to_string(i) + "]";
        return p + me;
    }
    return p.empty() ? me : p + "." + me;
}

Setting& Setting::getParent() const {
    if (!parent_) throw SettingNotFoundException("parent of root");
    return *parent_;
}

Setting& Setting::operator[](int idx) const {
    if (!isAggregate() || idx < 0 || idx >= (int)children_.size())
        throw SettingNotFoundException(getPath() + "[" + std::to_string(idx) + "]");
    return *children_[idx];
}

Setting& Setting::operator[](const char* name) const {
    if (isGroup())
        for (auto& c : children_)
            if (c->name_ == name) return *c;
    throw SettingNotFoundException(getPath().empty() ? name : getPath() + "." + name);
}

bool Setting::exists(const char* name) const {
    if (!isGroup()) return false;
    for (auto& c : children_)
        if (c->name_ == name) return true;
    return false;
}

int64_t Setting::as_int64() const {
    switch (type_) {
        case TypeInt:
        case TypeInt64:
            return ival_;
        case TypeFloat:
            return (int64_t)llround(fval_);
        case TypeBoolean:
            return bval_ ? 1 : 0;
        default:
            throw SettingTypeException(getPath());
    }
}

Setting::operator bool() const {
    if (type_ != TypeBoolean) throw SettingTypeException(getPath());
    return bval_;
}
Setting::operator int() const {
    if (type_ == TypeBoolean || !isNumber()) throw SettingTypeException(getPath());
    int64_t v = as_int64();
    if (type_ == TypeInt64 && (v > INT_MAX || v < INT_MIN)) throw SettingTypeException(getPath());
    return (int)v;
}
Setting::operator unsigned int() const {
    if (!isNumber()) throw SettingTypeException(getPath());
    return (unsigned int)as_int64();
}
Setting::operator long() const {
    if (!isNumber()) throw SettingTypeException(getPath());
    return (long)as_int64();
}
Setting::operator unsigned long() const {
    if (!isNumber()) throw SettingTypeException(getPath());
    return (unsigned long)as_int64();
}
Setting::operator long long() const {
    if (!isNumber()) throw SettingTypeException(getPath());
    return (long long)as_int64();
}
Setting::operator unsigned long long() const {
    if (!isNumber()) throw SettingTypeException(getPath());
    return (unsigned long long)as_int64();
}
Setting::operator double() const {
    if (type_ == TypeFloat) return fval_;
    if (type_ == TypeInt || type_ == TypeInt64) return (double)ival_;
    throw SettingTypeException(getPath());
}
Setting::operator float() const { return (float)(double)*this; }
Setting::operator std::string() const {
    if (type_ != TypeString) throw SettingTypeException(getPath());
    return sval_;
}
Setting::operator const char*() const {
    if (type_ != TypeString) throw SettingTypeException(getPath());
    return sval_.c_str();
}

bool Setting::lookupValue(const char* name, std::string& value) const {
    if (!isGroup() || !exists(name)) return false;
    const Setting& s = (*this)[name];
    if (s.type_ != TypeString) return false;
    value = s.sval_;
    return true;
}

void Setting::set_scalar_type(Type t) {
    if (isAggregate()) throw SettingTypeException(getPath());
    // Scalars inside an array must keep one type (libconfig rule).
    if (type_ != TypeNone && type_ != t) {
        const bool both_int = (type_ == TypeInt || type_ == TypeInt64) && (t == TypeInt || t == TypeInt64);
        if (!both_int) throw SettingTypeException(getPath());
    }
    type_ = t;
}

Setting& Setting::operator=(bool v) {
    set_scalar_type(TypeBoolean);
    bval_ = v;
    return *this;
}
Setting& Setting::operator=(int v) {
    if (type_ != TypeInt64) set_scalar_type(TypeInt);
    ival_ = v;
    return *this;
}
Setting& Setting::operator=(long v) { return *this = (long long)v; }
Setting& Setting::operator=(long long v) {
    if (type_ == TypeInt && v >= INT_MIN && v <= INT_MAX) {
        ival_ = v;
        return *this;
    }
    set_scalar_type(type_ == TypeInt ? TypeInt64 : (type_ == TypeNone ? TypeInt64 : type_));
    if (type_ != TypeInt && type_ != TypeInt64) throw SettingTypeException(getPath());
    ival_ = v;
    return *this;
}
Setting& Setting::operator=(double v) {
    set_scalar_type(TypeFloat);
    fval_ = v;
    return *this;
}
Setting& Setting::operator=(const char* v) {
    set_scalar_type(TypeString);
    sval_ = v ? v : "";
    return *this;
}

static bool valid_name(const char* n) {
    if (!n || !*n) return false;
    if (!(isalpha((unsigned char)n[0]) || n[0] == '*')) return false;
    for (const char* p = n + 1; *p; ++p)
        if (!(isalnum((unsigned char)*p) || *p == '_' || *p == '-' || *p == '*')) return false;
    return true;
}

Setting& Setting::add(const char* name, Type t) {
    if (!isGroup()) throw SettingTypeException(getPath());
    if (!valid_name(name) || exists(name)) throw SettingNameException(name ? name : "");
    children_.emplace_back(new Setting(this, name, t));
    return *children_.back();
}

Setting& Setting::add(Type t) {
    if (type_ != TypeList && type_ != TypeArray) throw SettingTypeException(getPath());
    if (type_ == TypeArray) {
        if (t == TypeGroup || t == TypeList || t == TypeArray) throw SettingTypeException(getPath());
        if (!children_.empty() && children_[0]->type_ != t &&
            !((children_[0]->type_ == TypeInt || children_[0]->type_ == TypeInt64) &&
              (t == TypeInt || t == TypeInt64)))
            throw SettingTypeException(getPath());
    }
    children_.emplace_back(new Setting(this, "", t));
    return *children_.back();
}

void Setting::remove(const char* name) {
    if (!isGroup()) throw SettingTypeException(getPath());
    for (size_t i = 0; i < children_.size(); ++i)
        if (children_[i]->name_ == name) {
            children_.erase(children_.begin() + i);
            return;
        }
    throw SettingNotFoundException(getPath() + "." + name);
}

void Setting::remove(unsigned int idx) {
    if (!isAggregate() || idx >= children_.size()) throw SettingNotFoundException(getPath());
    children_.erase(children_.begin() + idx);
}

static void write_string(std::string& out, const std::string& s) {
    out += '"';
    for (char c : s) {
        switch (c) {
            case '"': out += "\\\""; break;
            case '\\': out += "\\\\"; break;
            case '\n': out += "\\n"; break;
            case '\t': out += "\\t"; break;
            default: out += c;
        }
    }
    out += '"';
}

void Setting::write(std::string& out, int indent) const {
    std::string pad(indent * 2, ' ');
    switch (type_) {
        case TypeGroup: {
            if (parent_) out += "{\n";
            for (auto& c : children_) {
                out += pad + (parent_ ? "  " : "") + c->name_ + " = ";
                c->write(out, indent + (parent_ ? 1 : 0));
                out += ";\n";
            }
            if (parent_) out += pad + "}";
            break;
        }
        case TypeList:
        case TypeArray: {
            out += type_ == TypeList ? "( " : "[ ";
            for (size_t i = 0; i < children_.size(); ++i) {
                if (i) out += ", ";
                children_[i]->write(out, indent + 1);
            }
            out += type_ == TypeList ? " )" : " ]";
            break;
        }
        case TypeInt: out += std::to_string(ival_); break;
        case TypeInt64: out += std::to_string(ival_) + "L"; break;
        case TypeFloat: {
            std::ostringstream o;
            o.precision(17);
            o << fval_;
            std::string s = o.str();
            if (s.find_first_of(".eE") == std::string::npos) s += ".0";
            out += s;
            break;
        }
        case TypeBoolean: out += bval_ ? "true" : "false"; break;
        case TypeString: write_string(out, sval_); break;
        default: break;
    }
}

// ------------------------------------------------------------ parser
class Parser {
  public:
    explicit Parser(const char* t) : p_(t) {}

    void parse_settings(Setting& group, char terminator) {
        while (true) {
            skip();
            if (terminator ? *p_ == terminator : *p_ == 0) return;
            if (!*p_) throw ParseException(line_, "unexpected end of input");
            std::string name = parse_name();
            skip();
            if (*p_ != '=' && *p_ != ':') throw ParseException(line_, "expected '=' or ':' after " + name);
            ++p_;
            if (group.exists(name.c_str())) throw ParseException(line_, "duplicate setting " + name);
            Setting& s = *new_child(group, name);
            parse_value(s);
            skip();
            if (*p_ == ';' || *p_ == ',') ++p_;
        }
    }

  private:
    const char* p_;
    int line_ = 1;

    Setting* new_child(Setting& parent, const std::string& name) {
        parent.children_.emplace_back(new Setting(&parent, name, Setting::TypeNone));
        return parent.children_.back().get();
    }

    void skip() {
        while (*p_) {
            if (*p_ == '\n') {
                ++line_;
                ++p_;
            } else if (isspace((unsigned char)*p_)) {
                ++p_;
            } else if (*p_ == '#' || (p_[0] == '/' && p_[1] == '/')) {
                while (*p_ && *p_ != '\n') ++p_;
            } else if (p_[0] == '/' && p_[1] == '*') {
                p_ += 2;
                while (*p_ && !(p_[0] == '*' && p_[1] == '/')) {
                    if (*p_ == '\n') ++line_;
                    ++p_;
                }
                if (!*p_) throw ParseException(line_, "unterminated comment");
                p_ += 2;
            } else {
                break;
            }
        }
    }

    std::string parse_name() {
        const char* s = p_;
        if (!(isalpha((unsigned char)*p_) || *p_ == '*')) throw ParseException(line_, "expected a setting name");
        while (isalnum((unsigned char)*p_) || *p_ == '_' || *p_ == '-' || *p_ == '*') ++p_;
        return std::string(s, p_);
    }

    void parse_value(Setting& s) {
        skip();
        char c = *p_;
        if (c == '{') {
            ++p_;
            s.type_ = Setting::TypeGroup;
            parse_settings(s, '}');
            ++p_;
        } else if (c == '(' || c == '[') {
            const char close = c == '(' ? ')' : ']';
            s.type_ = c == '(' ? Setting::TypeList : Setting::TypeArray;
            ++p_;
            while (true) {
                skip();
                if (*p_ == close) {
                    ++p_;
                    break;
                }
                Setting* e = new_child(s, "");
                parse_value(*e);
                if (s.type_ == Setting::TypeArray) {
                    if (e->isAggregate()) throw ParseException(line_, "arrays hold scalars only");
                    const auto t0 = s.children_[0]->type_;
                    const bool ints = (t0 == Setting::TypeInt || t0 == Setting::TypeInt64) &&
                                      (e->type_ == Setting::TypeInt || e->type_ == Setting::TypeInt64);
                    if (e->type_ != t0 && !ints) throw ParseException(line_, "mixed types in array");
                }
                skip();
                if (*p_ == ',') {
                    ++p_;
                } else if (*p_ != close) {
                    throw ParseException(line_, "expected ',' or closing bracket");
                }
            }
        } else if (c == '"') {
            s.type_ = Setting::TypeString;
            s.sval_.clear();
            while (*p_ == '"') {
                ++p_;
                while (*p_ != '"') {
                    if (!*p_ || *p_ == '\n') throw ParseException(line_, "unterminated string");
                    if (*p_ == '\\') {
                        ++p_;
                        switch (*p_) {
                            case 'n': s.sval_ += '\n'; break;
                            case 't': s.sval_ += '\t'; break;
                            case 'r': s.sval_ += '\r'; break;
                            case 'f': s.sval_ += '\f'; break;
                            case '\\': s.sval_ += '\\'; break;
                            case '"': s.sval_ += '"'; break;
                            default: throw ParseException(line_, "bad escape");
                        }
                        ++p_;
                    } else {
                        s.sval_ += *p_++;
                    }
                }
                ++p_;
                skip();  // adjacent literals concatenate
            }
        } else if (isalpha((unsigned char)c)) {
            const char* st = p_;
            while (isalpha((unsigned char)*p_)) ++p_;
            std::string w(st, p_);
            for (auto& ch : w) ch = (char)tolower(ch);
            if (w == "true" || w == "false") {
                s.type_ = Setting::TypeBoolean;
                s.bval_ = w == "true";
            } else {
                throw ParseException(line_, "unexpected word " + w);
            }
        } else if (c == '-' || c == '+' || c == '.' || isdigit((unsigned char)c)) {
            parse_number(s);
        } else {
            throw ParseException(line_, std::string("unexpected character '") + c + "'");
        }
    }

    void parse_number(Setting& s) {
        const char* st = p_;
        bool neg = false;
        if (*p_ == '+' || *p_ == '-') neg = *p_++ == '-';
        if (p_[0] == '0' && (p_[1] == 'x' || p_[1] == 'X')) {
            char* end = nullptr;
            errno = 0;
            unsigned long long v = strtoull(p_ + 2, &end, 16);
            if (end == p_ + 2 || errno) throw ParseException(line_, "bad hex number");
            p_ = end;
            bool is64 = false;
            if (*p_ == 'L') {
                is64 = true;
                ++p_;
            }
            s.ival_ = neg ? -(long long)v : (long long)v;
            s.type_ = (is64 || v > 0xFFFFFFFFull) ? Setting::TypeInt64 : Setting::TypeInt;
            if (s.type_ == Setting::TypeInt) s.ival_ = (int32_t)(uint32_t)v * (neg ? -1 : 1);
            return;
        }
        const char* q = p_;
        while (isdigit((unsigned char)*q)) ++q;
        const bool is_float = *q == '.' || *q == 'e' || *q == 'E';
        if (is_float) {
            char* end = nullptr;
            double v = strtod(st, &end);
            if (end == st) throw ParseException(line_, "bad number");
            p_ = end;
            s.type_ = Setting::TypeFloat;
            s.fval_ = v;
            return;
        }
        char* end = nullptr;
        errno = 0;
        long long v = strtoll(st, &end, 10);
        if (end == st || errno) throw ParseException(line_, "bad integer");
        p_ = end;
        if (*p_ == 'L') {
            ++p_;
            s.type_ = Setting::TypeInt64;
        } else {
            if (v > INT_MAX || v < INT_MIN) throw ParseException(line_, "integer out of range (use the L suffix)");
            s.type_ = Setting::TypeInt;
        }
        s.ival_ = v;
    }
};

// ------------------------------------------------------------ Config
Config::Config() : root_(new Setting(nullptr, "", Setting::TypeGroup)) {}
Config::~Config() = default;

void Config::readString(const char* text) {
    root_.reset(new Setting(nullptr, "", Setting::TypeGroup));
    Parser p(text ? text : "");
    p.parse_settings(*root_, 0);
}

void Config::readFile(const char* filename) {
    std::ifstream f(filename);
    if (!f) throw ConfigException(std::string("cannot open ") + filename);
    std::stringstream ss;
    ss << f.rdbuf();
    readString(ss.str());
}

std::string Config::toString() const {
    std::string out;
    root_->write(out, 0);
    return out;
}

void Config::writeFile(const char* filename) const {
    std::ofstream f(filename);
    f << toString();
}

Setting& Config::lookup(const char* path) const {
    Setting* s = root_.get();
    std::string p(path);
    size_t i = 0;
    while (i < p.size()) {
        if (p[i] == '.') {
            ++i;
            continue;
        }
        if (p[i] == '[') {
            size_t j = p.find(']', i);
            if (j == std::string::npos) throw SettingNotFoundException(path);
            s = &(*s)[atoi(p.substr(i + 1, j - i - 1).c_str())];
            i = j + 1;
            continue;
        }
        size_t j = p.find_first_of(".[", i);
        std::string name = p.substr(i, j == std::string::npos ? std::string::npos : j - i);
        s = &(*s)[name.c_str()];
        i = j == std::string::npos ? p.size() : j;
    }
    return *s;
}

bool Config::exists(const char* path) const {
    try {
        lookup(path);
        return true;
    } catch (const SettingNotFoundException&) {
        return false;
    }
}

}  // namespace libconfig
