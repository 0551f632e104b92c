# The oracle as a deck runner (tlo_deck.c) under AddressSanitizer and
# UndefinedBehaviorSanitizer, host code only, leak detection on by default:
#   make -f oracle/sanitize.mk
#   oracle/tlo_deck_sanitized deck1.in deck2.in ...
# (with ASAN_OPTIONS=verify_asan_link_order=0 where another library is
# preloaded into every process).  About 8 s per 300x300 three-step deck.
ORACLE := $(patsubst %/,%,$(dir $(lastword $(MAKEFILE_LIST))))
CC_ORACLE ?= gcc

all: $(ORACLE)/tlo_deck_sanitized

$(ORACLE)/tlo_deck_sanitized: $(ORACLE)/tl_oracle.c $(ORACLE)/tl_oracle.h $(ORACLE)/tlo_deck.c
	$(CC_ORACLE) -std=c99 -O1 -g -fno-omit-frame-pointer -ffp-contract=off -Wall -Wextra \
	  -Wno-unused-parameter -fsanitize=address,undefined $(ORACLE)/tl_oracle.c \
	  $(ORACLE)/tlo_deck.c -o $@ -lm

clean:
	rm -f $(ORACLE)/tlo_deck_sanitized

.PHONY: all clean
