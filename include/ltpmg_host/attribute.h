// attribute.h -- one column of a table schema (the reference's
// include/attribute.h: same class and accessors).
#pragma once

#ifndef LTPMG_ATTRIBUTE
#define LTPMG_ATTRIBUTE

#include "define.h"

class Attribute {
public:
    Attribute(const std::string &name, const std::string &type, const std::string &data_distribution)
        : name(name), type(type), data_distribution(data_distribution) {}

    const std::string &getName() const;

    const std::string &getType() const;

private:
    std::string name;               // column name
    std::string type;               // "int", "double" or "string"
    std::string data_distribution;  // kept as read; the fill is seeded (include/ltpmg.h)
};

#endif
