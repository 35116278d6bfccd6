#pragma once
// the rank is kept in `level`, as in the reference application
#define MEDUSA_VERTEX_FIELDS(F) F(MVT, level)
