#pragma once
#define MEDUSA_VERTEX_FIELDS(F) F(MVT, frontier) F(MVT, visited) F(MVT, next)
