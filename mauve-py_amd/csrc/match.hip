// match.hip -- placeholder
#include "mh_internal.hpp"
