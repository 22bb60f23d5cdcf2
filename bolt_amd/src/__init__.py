"""Built-in physics plugins (torch mirrors of bolt/src/*)."""
