#pragma once
#define MEDUSA_EDGE_FIELDS(F)
