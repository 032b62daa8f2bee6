/* bluebottle executable: the whole program of src/bluebottle.c:28-436,
 * run from a directory that holds input/flow.config and input/decomp.config. */
#include "bluebottle.h"

int main(int argc, char *argv[])
{
  return bluebottle_main(argc, argv);
}
