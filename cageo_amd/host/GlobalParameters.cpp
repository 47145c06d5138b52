// Process-wide parameters read by the stages (GlobalParameters.cpp of the
// reference: noDataValueDEM and noDataValueStream default to 0 and are set
// from the raster headers by the driver).
#include "Stages.h"

namespace cageo_host {
namespace G {
double noDataValueDEM = 0.0;
double noDataValueStream = 0.0;
}  // namespace G

AbstractFactory::Type AbstractFactory::defaultImpl = AbstractFactory::GPU;

}  // namespace cageo_host
