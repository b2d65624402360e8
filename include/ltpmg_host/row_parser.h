// row_parser.h -- reads a schema file (the reference's include/row_parser.h:
// same class; the format is the reference's data/<benchmark>.json:
// {"tables": [{"name", "size", "attributes": [{"name", "type",
// "data_distribution"}]}]}).  A malformed file is a fatal error, as a
// failed parse is in the reference.
#pragma once

#ifndef LTPMG_ROW_PARSER
#define LTPMG_ROW_PARSER

#include "row.h"

class Row_Parser {
public:
    const std::vector<Row> &parse(const std::string &filename);

private:
    std::vector<Row> tables;
};

#endif
