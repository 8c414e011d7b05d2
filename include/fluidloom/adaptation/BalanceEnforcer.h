// BalanceEnforcer (the reference declares each engine in its own header).
#pragma once
#include "fluidloom/adaptation/Engines.h"
