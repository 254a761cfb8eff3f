#include "schema.h"

#include <algorithm>
#include <sstream>

void Schema::add(ColumnType ct, unsigned int size) {
    unsigned int s = 0;
    switch (ct) {
        case CT_INTEGER: s = sizeof(CtInt); break;
        case CT_LONG: s = sizeof(CtLong); break;
        case CT_DECIMAL: s = sizeof(CtDecimal); break;
        case CT_DATE: s = sizeof(CtDate); break;
        case CT_POINTER: s = sizeof(void*); break;
        case CT_CHAR:
            if (size < 2) throw IllegalSchemaDeclarationException();
            s = size;
            break;
        default: throw IllegalSchemaDeclarationException();
    }
    vct_.push_back(ct);
    voffset_.push_back(totalsize_);
    vsize_.push_back(s);
    vfmt_.emplace_back();
    totalsize_ += s;
}

void Schema::add(ColumnSpec cs) {
    add(cs.type, cs.type == CT_CHAR ? cs.size : 0);
    vfmt_.back() = cs.formatstr;
}

ColumnSpec Schema::get(unsigned int pos) const { return ColumnSpec(vct_.at(pos), vsize_.at(pos), vfmt_.at(pos)); }

void Schema::writeData(void* dest, unsigned int pos, const void* data) const {
    char* d = (char*)dest + voffset_[pos];
    if (vct_[pos] == CT_CHAR) {
        strncpy(d, (const char*)data, vsize_[pos]);
    } else {
        memcpy(d, data, vsize_[pos]);
    }
}

CtInt Schema::asInt(const void* tuple, unsigned int pos) const {
    if (vct_[pos] != CT_INTEGER) throw IllegalConversionException();
    CtInt v;
    memcpy(&v, (const char*)tuple + voffset_[pos], sizeof v);
    return v;
}

CtLong Schema::asLong(const void* tuple, unsigned int pos) const {
    const char* p = (const char*)tuple + voffset_[pos];
    switch (vct_[pos]) {
        case CT_INTEGER: {
            CtInt v;
            memcpy(&v, p, sizeof v);
            return v;
        }
        case CT_LONG:
        case CT_DATE: {
            CtLong v;
            memcpy(&v, p, sizeof v);
            return v;
        }
        default: throw IllegalConversionException();
    }
}

CtDecimal Schema::asDecimal(const void* tuple, unsigned int pos) const {
    const char* p = (const char*)tuple + voffset_[pos];
    switch (vct_[pos]) {
        case CT_DECIMAL: {
            CtDecimal v;
            memcpy(&v, p, sizeof v);
            return v;
        }
        case CT_INTEGER:
        case CT_LONG:
            return (CtDecimal)asLong(tuple, pos);
        default: throw IllegalConversionException();
    }
}

std::string Schema::prettyprint(const void* tuple, char sep) const {
    std::ostringstream o;
    for (unsigned int i = 0; i < columns(); ++i) {
        if (i) o << sep;
        switch (vct_[i]) {
            case CT_INTEGER: o << asInt(tuple, i); break;
            case CT_LONG:
            case CT_DATE: o << asLong(tuple, i); break;
            case CT_DECIMAL: o << asDecimal(tuple, i); break;
            case CT_CHAR: o << std::string((const char*)tuple + voffset_[i], strnlen((const char*)tuple + voffset_[i], vsize_[i])); break;
            default: o << "?";
        }
    }
    return o.str();
}

Schema Schema::create(const libconfig::Setting& line) {
    Schema ret;
    for (int i = 0; i < line.getLength(); ++i) {
        std::string val = line[i];
        size_t paren = val.find('(');
        std::transform(val.begin(), paren == std::string::npos ? val.end() : val.begin() + paren, val.begin(),
                       ::tolower);
        if (val.rfind("int", 0) == 0) {
            ret.add(CT_INTEGER);
        } else if (val.rfind("long", 0) == 0) {
            ret.add(CT_LONG);
        } else if (val.rfind("dec", 0) == 0) {
            ret.add(CT_DECIMAL);
        } else if (val.rfind("char", 0) == 0) {
            if (paren == std::string::npos) throw IllegalSchemaDeclarationException();
            int len = atoi(val.c_str() + paren + 1);
            ret.add(CT_CHAR, len + 1);  // room for the terminating zero
        } else if (val.rfind("date", 0) == 0) {
            if (paren == std::string::npos) throw IllegalSchemaDeclarationException();
            ret.add(ColumnSpec(CT_DATE, 8, val.substr(paren + 1)));
        } else {
            throw IllegalSchemaDeclarationException();
        }
    }
    return ret;
}
