#include "common.hpp"
