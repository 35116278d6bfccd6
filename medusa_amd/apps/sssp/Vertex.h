#pragma once
#define MEDUSA_VERTEX_FIELDS(F) F(MVT, distance) F(bool, updated)
