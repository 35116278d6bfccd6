#pragma once
#define MVT float
#define MVT_Init_Value -1
