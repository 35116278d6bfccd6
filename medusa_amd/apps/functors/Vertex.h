#pragma once
#define MEDUSA_VERTEX_FIELDS(F) F(MVT, wsum) F(MVT, result)
