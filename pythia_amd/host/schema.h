// Packed-row schema with the reference's column types and layout rules:
// column offsets are the running sum of widths with no alignment padding
// (schema.cpp:82-117); int = 4 B, long / dec / date = 8 B, char(n) = n+1 B.
#pragma once

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "plan_config.h"

enum ColumnType { CT_INTEGER, CT_LONG, CT_DECIMAL, CT_CHAR, CT_DATE, CT_POINTER };

typedef int CtInt;
typedef long long CtLong;
typedef double CtDecimal;
typedef char CtChar;
typedef long long CtDate;

struct ColumnSpec {
    ColumnType type;
    unsigned int size;
    std::string formatstr;  // date format, unused by the hot path
    ColumnSpec(ColumnType t = CT_INTEGER, unsigned int s = 4, std::string f = "") : type(t), size(s), formatstr(std::move(f)) {}
};

class IllegalSchemaDeclarationException : public std::runtime_error {
  public:
    IllegalSchemaDeclarationException() : std::runtime_error("illegal schema declaration") {}
};
class IllegalConversionException : public std::runtime_error {
  public:
    IllegalConversionException() : std::runtime_error("illegal conversion") {}
};

class Schema {
  public:
    Schema() = default;
    void add(ColumnType ct, unsigned int size = 0);
    void add(ColumnSpec cs);
    unsigned int columns() const { return (unsigned int)vct_.size(); }
    ColumnSpec get(unsigned int pos) const;
    ColumnType getColumnType(unsigned int pos) const { return vct_.at(pos); }
    unsigned int getColumnWidth(unsigned int pos) const { return vsize_.at(pos); }
    unsigned int getColumnOffset(unsigned int pos) const { return voffset_.at(pos); }
    unsigned int getTupleSize() const { return totalsize_; }

    void* calcOffset(void* tuple, unsigned int pos) const { return (char*)tuple + voffset_[pos]; }
    const void* calcOffset(const void* tuple, unsigned int pos) const { return (const char*)tuple + voffset_[pos]; }
    void writeData(void* dest, unsigned int pos, const void* data) const;
    void copyTuple(void* dest, const void* src) const { memcpy(dest, src, totalsize_); }

    CtInt asInt(const void* tuple, unsigned int pos) const;
    CtLong asLong(const void* tuple, unsigned int pos) const;
    CtDecimal asDecimal(const void* tuple, unsigned int pos) const;
    std::string prettyprint(const void* tuple, char sep) const;

    // "int", "long", "dec", "char(n)", "date(fmt)" (schema.cpp:221-259).
    static Schema create(const libconfig::Setting& line);

  private:
    std::vector<ColumnType> vct_;
    std::vector<unsigned int> voffset_, vsize_;
    std::vector<std::string> vfmt_;
    unsigned int totalsize_ = 0;
};
