// OPM_THROW for code emitted by the Equelle compiler (the reference gets it from
// opm-common; this runtime ships its own definition with the same meaning:
// throw `exception_type` with a message built from a stream expression).
#ifndef EQUELLE_AMD_OPM_ERRORMACROS_HPP
#define EQUELLE_AMD_OPM_ERRORMACROS_HPP
#include <sstream>
#include <stdexcept>

#define OPM_THROW(Exception, message)              \
    do {                                           \
        std::ostringstream opm_throw_os_;          \
        opm_throw_os_ << message;                  \
        throw Exception(opm_throw_os_.str());      \
    } while (false)

#endif
